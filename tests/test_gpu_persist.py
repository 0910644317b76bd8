"""Batched shots in one persistent launch per run: every step of every shot from one queue of
work items, shots ordered only by their own item counters. The per-cell arithmetic does not
depend on which workgroup runs a tile or when, so records, wavefields, PML state and per-step
maxima must equal each shot's plan run alone bit for bit: with every tile set, both ways of
cutting the items, runs split at an odd step (slot parity carried across launches), a source on
the frozen rim, shots with and without compact y-face state, and 16-bit storage. A wait bound
forced to zero from the host makes the launch leave early and the sync report it (every
workgroup still exits normally); the next run is exact again."""
import sys

import numpy as np
import pytest

import qp3d

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def shots(nt, storage="fp32", rim=False, mixed=False):
    model = qp3d.genmodel(70, 64, 30, 3)
    if mixed:
        cases = [((0, 0, 31, 27), 10), ((20, 9, 24, 20), 34), ((5, 30, 26, 23), 32)]
    else:
        cases = [((0, 0, 31, 27), 10), ((20, 9, 44, 40), 10), ((35, 30, 26, 33), 10), ((5, 40, 60, 21), 10)]
    setups = [qp3d.window_setup(model, x0, y0, hx, hy, pml=pml, nt=nt, fm=20.0, src_depth=2)
              for (x0, y0, hx, hy), pml in cases]
    if rim:
        s = setups[1]
        src = np.array(s.source, copy=True)
        src[2, s.nx // 2, s.ny // 2 + 1] = 0.25  # on the frozen rim (z < 5)
        s.source = src
    return [qp3d.Plan(s, storage=storage) for s in setups]


def state(p):
    return (p.record(), p.field(0), p.field(1), p.field(2), p.field(4), p.field(6), p.field(8), p.step_maxima())


def alone(plans):
    out = []
    for p in plans:
        p.reset()
        p.run()
        p.sync()
        out.append(state(p))
    return out


def together(b, plans, split=7):
    b.reset()
    b.run(0, split)
    b.run(split, plans[0].s.nt)
    b.sync()
    return [state(p) for p in plans]


def same(a, b):
    for i, (x, y) in enumerate(zip(a, b)):
        for j, (u, v) in enumerate(zip(x, y)):
            assert np.array_equal(u, v), (i, j, float(np.abs(u - v).max()))


@pytest.mark.parametrize("tile_set", [0, 1, 2])
def test_persistent_equals_shots_alone(tile_set):
    plans = shots(nt=40, rim=True)
    try:
        ref = alone(plans)
        with qp3d.Batch(plans) as b:
            for share in (False, True):
                say(f"persistent fp32 tile set {tile_set} share {share}")
                b.set_persistent(tile_set, share)
                assert b.persistent() == tile_set
                same(together(b, plans, split=7), ref)
                same(together(b, plans, split=22), ref)
    finally:
        for p in plans:
            p.close()


@pytest.mark.parametrize("storage", ["bf16", "fp16"])
def test_persistent_16bit_storage(storage):
    plans = shots(nt=30, storage=storage)
    try:
        ref = alone(plans)
        with qp3d.Batch(plans) as b:
            say(f"persistent {storage}")
            b.set_persistent(0, False)
            same(together(b, plans), ref)
    finally:
        for p in plans:
            p.close()


def test_persistent_mixes_compact_and_full_grid_y_faces():
    plans = shots(nt=30, mixed=True)
    try:
        ref = alone(plans)
        with qp3d.Batch(plans) as b:
            for tile_set in (0, 2):
                say(f"persistent mixed pml, tile set {tile_set}")
                b.set_persistent(tile_set, True)
                same(together(b, plans), ref)
    finally:
        for p in plans:
            p.close()


def test_persistent_wait_bound_reports_and_recovers():
    plans = shots(nt=30)
    try:
        ref = alone(plans)
        with qp3d.Batch(plans) as b:
            b.set_persistent(0, False)
            b.set_wait_limit(0.0)
            say("persistent with a zero wait bound")
            b.reset()
            b.run()
            with pytest.raises(qp3d.Qp3dError) as ei:
                b.sync()
            assert ei.value.code == -7
            b.set_wait_limit(5e6)
            say("persistent after the bound is restored")
            same(together(b, plans), ref)
    finally:
        for p in plans:
            p.close()


def test_autotune_considers_the_persistent_launch():
    plans = shots(nt=30)
    try:
        ref = alone(plans)
        with qp3d.Batch(plans) as b:
            t = b.autotune(steps=3)
            say(f"autotuned {t}")
            assert t["persistent"] in (-1, 0, 1, 2)
            same(together(b, plans), ref)
            b.set_persistent(1, True)
            same(together(b, plans), ref)
            b.set_persistent(-1, False)
            assert b.persistent() == -1
            same(together(b, plans), ref)
    finally:
        for p in plans:
            p.close()
