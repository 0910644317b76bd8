"""Amplitude bound of the gradient pass: the pass skips tiles whose bound 2 sum|c1|/d max|u|
(from the block maxima the updates record) is below the G already found exactly, so G_k and
every result must be the same with the bound on and off, bit for bit: direct launches and the
captured graph, 16-bit storage, a source on the frozen rim, and a plan whose block maxima went
stale while a batch drove it. The bound must also actually skip tiles on a shot from rest.

The same maxima drive the quiet map (S = 1 proved: no eps - del load) and the rest map (blocks
the wave has not reached, left out of the updates): both are covered by the same comparisons,
and a shot whose rest history breaks halfway (steps run without the map) must stay exact."""
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


@pytest.mark.parametrize("persistent", [False, True])
def test_bound_after_a_batch_drove_the_plan(persistent):
    """A batch (launches per stage, or the persistent launch) leaves the plan's block maxima
    stale and its rest history broken: the plan's next run from the batch's state must evaluate
    every tile of its first step, count every block as touched, and match a run without the
    bound."""
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
                if persistent:
                    b.set_persistent(0, False)
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


@pytest.mark.parametrize("graph", [False, True])
def test_rest_map_leaves_out_cells_and_changes_nothing(graph):
    s = qp3d.config_setup(hx=101, hy=101, nz0=101, pml=20, nt=300, seed=0, src_depth=3)
    with qp3d.Plan(s, use_graph=graph) as p:
        p.set_amp_bound(False)
        ref = run(p)
        p.set_amp_bound(True)
        p.set_instrument(1)
        say("rest map on, instrumented")
        same(run(p), ref)
        uc = p.update_cells()
        p.set_instrument(0)
        say(f"update cells per kind (cells, at rest, quiet): {uc.tolist()}")
        # every cell of [5, n-5)^3 is in exactly one update item per step
        assert int(uc[:, 0].sum()) == s.nt * (s.nx - 10) * (s.ny - 10) * (s.nz - 10)
        assert uc[0, 1] > 0.05 * uc[0, 0] and uc[1, 1] + uc[2, 1] > 0.05 * (uc[1, 0] + uc[2, 0])
        assert uc[0, 2] > 0
        assert np.all(uc[:, 1] + uc[:, 2] <= uc[:, 0])
        # the history breaks for ten steps (bound off), then the map counts every block as touched
        p.reset()
        p.run(0, 31)
        p.set_amp_bound(False)
        p.run(31, 41)
        p.set_amp_bound(True)
        p.run(41, s.nt)
        p.sync()
        say(f"rest history broken at step 31, bound back on at 41, graph={graph}")
        same(state(p), ref)


def test_rest_map_with_a_source_deep_in_the_grid():
    """A source far from every face: most blocks stay at rest for many steps, and the first to
    leave rest sit next to the source box."""
    s = qp3d.config_setup(hx=61, hy=67, nz0=59, pml=12, nt=160, seed=3, src_depth=0)
    src = np.zeros_like(s.source)
    src[s.nz // 2, s.nx // 2 + 3, s.ny // 2 - 2] = 1.0
    src[s.nz // 2 + 9, s.nx // 2 - 8, s.ny // 2 + 7] = -0.7  # another block
    s.source = src
    with qp3d.Plan(s) as p:
        p.set_amp_bound(False)
        ref = run(p)
        p.set_amp_bound(True)
        p.set_instrument(1)
        same(run(p, split=23), ref)
        uc = p.update_cells()
        p.set_instrument(0)
        say(f"deep source: update cells per kind {uc.tolist()}")
        # small grid: the touched region (whole 8 x 8 x 32 blocks and their neighbours) covers
        # it within some tens of steps
        assert uc[:, 1].sum() > 0.005 * uc[:, 0].sum()


@pytest.mark.parametrize("graph", [False, True])
def test_two_streams_equal_one(graph):
    """PML updates on a second stream, concurrent with the interior update (in a captured graph:
    forked and joined inside it), with the bound and both maps on: the same shot bit for bit."""
    s = qp3d.config_setup(hx=101, hy=101, nz0=101, pml=20, nt=300, seed=0, src_depth=3)
    with qp3d.Plan(s) as p:
        ref = run(p)
    with qp3d.Plan(s, use_graph=graph, streams=2) as p:
        say(f"two streams, graph={graph}")
        same(run(p, split=37), ref)
        same(run(p), ref)
