"""Batched shots: one launch per stage for several shots with uneven windows must give the
same records, state and per-step maxima as running each shot's plan alone (bit for bit: the
per-cell arithmetic does not depend on how cells are grouped into work items), with direct
launches, with the captured graph, with each shot's own launches, after the autotuner, and
with 16-bit storage."""
import sys

import numpy as np
import pytest

import qp3d

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def survey(nt=50, storage="fp32", gmax_mode=1):
    model = qp3d.genmodel(70, 64, 30, 3)
    wins = [(0, 0, 31, 27), (20, 9, 44, 40), (35, 30, 26, 33), (5, 40, 60, 21)]
    setups = [qp3d.window_setup(model, x0, y0, hx, hy, pml=10, nt=nt, fm=20.0, src_depth=2)
              for x0, y0, hx, hy in wins]
    return setups, [qp3d.Plan(s, storage=storage, gmax_mode=gmax_mode) for s in setups]


def alone(plans):
    out = []
    for p in plans:
        p.reset()
        p.run()
        p.sync()
        out.append((p.record(), p.field(0), p.field(2), p.step_maxima()))
    return out


def together(b, plans):
    b.reset()
    b.run(0, 7)
    b.run(7, plans[0].s.nt)
    b.sync()
    return [(p.record(), p.field(0), p.field(2), p.step_maxima()) for p in plans]


def same(a, b):
    for x, y in zip(a, b):
        for u, v in zip(x, y):
            assert np.array_equal(u, v)


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_batch_equals_shots_alone(storage):
    setups, plans = survey(storage=storage)
    try:
        ref = alone(plans)
        for graph in (False, True):
            say(f"batch {storage} graph={graph}")
            with qp3d.Batch(plans, use_graph=graph) as b:
                same(together(b, plans), ref)
                b.set_per_shot(True)
                same(together(b, plans), ref)
                b.set_per_shot(False)
                if not graph:
                    t = b.autotune(steps=3)
                    say(f"autotuned {t}")
                    assert all(1 <= x <= qp3d.NTILES for x in t["tiles"])
                    same(together(b, plans), ref)
    finally:
        for p in plans:
            p.close()


def test_batch_mixes_compact_and_full_grid_y_faces():
    """Shots whose y faces keep their PML state in compact arrays (pml <= 32, folded faces) and a
    shot whose do not (pml 34) in one batch: the y-face stage launches both update kinds, and
    every shot matches its plan run alone."""
    model = qp3d.genmodel(70, 64, 30, 3)
    cases = [((0, 0, 31, 27), 10), ((20, 9, 24, 20), 34), ((5, 30, 26, 23), 32)]
    setups = [qp3d.window_setup(model, x0, y0, hx, hy, pml=pml, nt=30, fm=20.0, src_depth=2)
              for (x0, y0, hx, hy), pml in cases]
    plans = [qp3d.Plan(s) for s in setups]
    try:
        ref = alone(plans)
        for graph in (False, True):
            say(f"mixed pml batch graph={graph}")
            with qp3d.Batch(plans, use_graph=graph) as b:
                same(together(b, plans), ref)
    finally:
        for p in plans:
            p.close()


def test_batch_rejects_a_plan_uploaded_again():
    setups, plans = survey(nt=12)
    try:
        with qp3d.Batch(plans) as b:
            plans[1].upload(setups[1])
            with pytest.raises(qp3d.Qp3dError):
                b.run()
    finally:
        for p in plans:
            p.close()


def test_batch_runs_predicted_threshold_plans_with_the_gradient_pass():
    """Plans created with gmax_mode 2 run inside a batch with the gradient pass (the batch
    kernels take G from it), bit-identical to gmax_mode 1 plans alone, and run with the
    predicted threshold again, bit-identical to before, once the batch is destroyed."""
    setups, plans1 = survey(nt=30)
    plans2 = [qp3d.Plan(s, gmax_mode=2) for s in setups]
    try:
        ref1, ref2 = alone(plans1), alone(plans2)
        for per_shot in (False, True):
            say(f"batch of gmax_mode 2 plans, per_shot={per_shot}")
            with qp3d.Batch(plans2) as b:
                b.set_per_shot(per_shot)
                same(together(b, plans2), ref1)
        same(alone(plans2), ref2)
    finally:
        for p in plans1 + plans2:
            p.close()


def test_batch_rejects_duplicate_and_shared_plans():
    """A plan holds one shot's wavefields, record and maxima: it cannot appear twice in a batch
    or belong to two live batches; once its batch is destroyed it can join another."""
    setups, plans = survey(nt=12)
    try:
        with pytest.raises(qp3d.Qp3dError) as ei:
            qp3d.Batch([plans[0], plans[1], plans[0]])
        assert ei.value.code == -1
        with qp3d.Batch(plans[:2]) as b:
            with pytest.raises(qp3d.Qp3dError) as ei:
                qp3d.Batch(plans[1:3])
            assert ei.value.code == -6
            b.forward()
        with qp3d.Batch(plans[1:3]) as b2:
            recs = b2.forward()
        assert np.array_equal(recs[0], alone([plans[1]])[0][0])
    finally:
        for p in plans:
            p.close()
