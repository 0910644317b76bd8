"""GPU parity: the HIP path through the C ABI against the CPU oracle (FP32, the
reference's own arithmetic) on the same seeded inputs. Tolerance: relative L2 <= 1e-5
on gathers, snapshots and state (the stated bar for FP32)."""
import os
import sys

import numpy as np
import pytest

from conftest import predict_G, rel_l2, s_fields_agree, steps_in_band

import qp3d
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_both(s, steps=None, predictions=None, **plan_kw):
    steps = s.nt if steps is None else steps
    say(f"case nx={s.nx} ny={s.ny} nz={s.nz} pml={s.pml} nt={steps} {plan_kw}")
    with qp3d.Plan(s, **plan_kw) as p:
        p.reset()
        p.run(0, steps)
        p.sync()
        rec = p.record()[:steps]
        st = {w: p.field(w) for w in range(9)}
        gm = p.step_maxima()[:steps]
        if predictions is not None:
            predictions.append(p.step_predictions()[:steps])
    o = O.OracleState(s)
    orec = np.zeros((steps, s.hx, s.hy), np.float32)
    og = np.zeros(steps)
    for k in range(steps):
        o.step(s.wavelet[k])
        og[k] = o.G
        orec[k] = o.field("u2")[s.sz, s.pml:s.nx - s.pml, s.pml:s.ny - s.pml]
    return rec, st, gm, orec, o, og


def test_small_asymmetric_all_pml_faces():
    s = qp3d.config_setup(hx=21, hy=27, nz0=19, pml=10, nt=60, seed=0, src_depth=3)
    rec, st, gm, orec, o, og = run_both(s)
    assert rel_l2(rec, orec) <= TOL
    assert rel_l2(st[0], o.field("u2")) <= TOL
    assert rel_l2(st[1], o.field("u1")) <= TOL
    np.testing.assert_allclose(gm, og, rtol=1e-5, atol=0)
    # the frozen 5-cell rim stays zero
    u = st[0]
    for sl in (u[:5], u[-5:], u[:, :5], u[:, -5:], u[:, :, :5], u[:, :, -5:]):
        assert not np.any(sl)


def _unified(o, s, arr):
    """Oracle per-face PML arrays scattered into the full-grid layout of the HIP state
    (each cell belongs to the first matching face: x, then y, then z)."""
    nx, ny, nz, pml = s.nx, s.ny, s.nz, s.pml
    out = np.zeros((nz, nx, ny), np.float64)
    z, x, y = np.meshgrid(np.arange(nz), np.arange(nx), np.arange(ny), indexing="ij")
    r5 = (x >= 5) & (x < nx - 5) & (y >= 5) & (y < ny - 5) & (z >= 5) & (z < nz - 5)
    xin = (x >= pml) & (x < nx - pml)
    yin = (y >= pml) & (y < ny - pml)
    zin = (z >= pml) & (z < nz - pml)
    faces = [
        ("L", r5 & (x < pml), z * ny * pml + x * ny + y),
        ("R", r5 & (x >= nx - pml), z * ny * pml + (x - nx + pml) * ny + y),
        ("F", r5 & xin & (y < pml), z * nx * pml + x * pml + y),
        ("Bk", r5 & xin & (y >= ny - pml), z * nx * pml + x * pml + (y - ny + pml)),
        ("T", r5 & xin & yin & (z < pml), z * nx * ny + x * ny + y),
        ("Bo", r5 & xin & yin & (z >= nz - pml), (z - nz + pml) * nx * ny + x * ny + y),
    ]
    for face, m, idx in faces:
        out[m] = o.aux(face, arr)[idx[m]]
    return out


def test_pml_state_matches_oracle_faces():
    s = qp3d.config_setup(hx=17, hy=15, nz0=13, pml=8, nt=40, seed=3, src_depth=2)
    rec, st, gm, orec, o, og = run_both(s)
    # newest slots hold w13/w33/p3 after the rotation, i.e. the oracle's w12/w32/p2
    pairs = [(2, "w12"), (3, "w11"), (4, "w32"), (5, "w31"), (6, "p2"), (7, "p1"), (8, "w21")]
    for which, arr in pairs:
        ref = _unified(o, s, arr)
        assert rel_l2(st[which], ref) <= TOL, (which, arr)


def test_source_on_frozen_rim():
    """A source with support outside [5,n-5)^3 accumulates there as in the reference."""
    s = qp3d.config_setup(hx=15, hy=15, nz0=15, pml=8, nt=30, seed=1, src_depth=0)
    src = np.zeros_like(s.source)
    src[2, 12, 14] = 1.0      # in the rim (z < 5)
    src[20, 3, 9] = -0.5      # in the rim (x < 5)
    src[12, 15, 15] = 0.75    # regular cell
    s.source = src
    rec, st, gm, orec, o, og = run_both(s)
    assert rel_l2(st[0], o.field("u2")) <= TOL
    assert rel_l2(rec, orec) <= TOL


def test_c1_geometry_reduced_steps():
    """The 101^3-interior, pml-20 grid of the CPU-runnable configuration, 300 steps."""
    s = qp3d.config_setup(hx=101, hy=101, nz0=101, pml=20, nt=300, seed=0, src_depth=3)
    say("c1 reduced: gpu")
    with qp3d.Plan(s) as p:
        rec = p.forward()
    say("c1 reduced: oracle")
    orec = O.modeling3d(s, "f32")
    err = rel_l2(rec, orec)
    say(f"c1 reduced rel-L2 {err:.3e}")
    assert err <= TOL


def test_dropin_modeling3d_snapshots(tmp_path):
    s = qp3d.config_setup(hx=23, hy=19, nz0=21, pml=10, nt=45, seed=2, src_depth=1)
    g = str(tmp_path / "gpu_snap")
    c = str(tmp_path / "cpu_snap")
    rec = qp3d.modeling3d(s, snapflag=1, ntsnap=20, snap_file=g)
    orec = O.modeling3d(s, "f32", snapflag=1, ntsnap=20, snap_file=c)
    assert rel_l2(rec, orec) <= TOL
    for k in (0, 20, 40):
        for suf in ("", "_xz", "_yz", "_xy"):
            a = np.fromfile(f"{g}{suf}{k}.dat", np.float32)
            b = np.fromfile(f"{c}{suf}{k}.dat", np.float32)
            assert a.size == b.size and a.size > 0
            if np.any(b):
                assert rel_l2(a, b) <= TOL, (k, suf)
            else:
                assert not np.any(a)
    assert not os.path.exists(f"{g}60.dat")


def test_source_scaling_is_exact_homogeneity():
    """The operator is positively homogeneous (S depends on grad u / |grad u| and on a
    threshold relative to max |grad u|), so doubling the source doubles every output."""
    s = qp3d.config_setup(hx=41, hy=37, nz0=33, pml=12, nt=150, seed=4, src_depth=2)
    with qp3d.Plan(s) as p:
        r1 = p.forward()
        s.source = s.source * np.float32(2)
        p.upload(s)
        r2 = p.forward()
    assert rel_l2(r2, 2 * r1) <= 1e-6


@pytest.mark.parametrize("gmax_mode", [1, 2])
def test_blowup_is_reported(gmax_mode):
    s = qp3d.config_setup(hx=21, hy=21, nz0=21, pml=8, nt=400, seed=0, dt=0.004)
    say(f"blow-up, gmax_mode {gmax_mode}")
    with qp3d.Plan(s, gmax_mode=gmax_mode) as p:
        p.reset()
        p.run()
        with pytest.raises(qp3d.Qp3dError) as ei:
            p.sync()
    assert ei.value.code == -5


@pytest.mark.parametrize("tile", list(range(1, 8)))
def test_every_tile_shape(tile):
    """Each launchable tile shape of the three stencil kernels against the oracle, on a grid
    whose extents are not multiples of any tile and with z-chunks shorter than the grid; odd
    shapes run the two updates on one stream, even shapes concurrently on two."""
    s = qp3d.config_setup(hx=37, hy=71, nz0=29, pml=9, nt=50, seed=7, src_depth=4)
    streams = 1 if tile % 2 else 2
    say(f"tile {tile} streams {streams}")
    with qp3d.Plan(s, tile_update=tile, tile_grad=tile, tile_pml=tile, tile_pml_y=8 - tile, zchunk=11,
                   streams=streams) as p:
        rec = p.forward()
        u = p.field(0)
    o = O.OracleState(s)
    orec = np.zeros_like(rec)
    for k in range(s.nt):
        o.step(s.wavelet[k])
        orec[k] = o.field("u2")[s.sz, s.pml:s.nx - s.pml, s.pml:s.ny - s.pml]
    assert rel_l2(rec, orec) <= TOL
    assert rel_l2(u, o.field("u2")) <= TOL


@pytest.mark.parametrize("pml", [6, 13, 32, 36])
def test_pml_widths_and_source_in_pml(pml):
    """PML widths around the folded y-face tiling limit (faces up to 27 cells are folded into
    one wavefront), with source cells inside the y-face PML strips."""
    s = qp3d.config_setup(hx=19, hy=23, nz0=17, pml=pml, nt=40, seed=8, src_depth=2)
    src = s.source.copy()
    src[pml + 3, s.nx // 2, 7] = 0.5          # low y face
    src[pml + 4, s.nx // 2 + 1, s.ny - 8] = -0.4  # high y face
    src[pml + 2, 7, s.ny // 2] = 0.3          # low x face
    s.source = src
    rec, st, gm, orec, o, og = run_both(s)
    assert rel_l2(rec, orec) <= TOL
    assert rel_l2(st[0], o.field("u2")) <= TOL
    for which, arr in [(2, "w12"), (6, "p2"), (8, "w21")]:
        assert rel_l2(st[which], _unified(o, s, arr)) <= TOL, (which, arr)


def test_graph_replay_matches_direct_launches():
    """Steps replayed from the captured 10-step graph give bit-identical records, state,
    per-step maxima and snapshots to direct launches, over runs split at odd steps, groups
    broken by snapshot steps, a source on the frozen rim, and a second shot replaying the
    same graph after host synchronisation."""
    s = qp3d.config_setup(hx=25, hy=30, nz0=22, pml=8, nt=57, seed=5, src_depth=2)
    src = s.source.copy()
    src[2, s.nx // 2, s.ny // 2 + 1] = 0.25  # on the frozen rim
    s.source = src
    say("graph: direct")
    with qp3d.Plan(s, snapflag=1, ntsnap=13) as p:
        p.reset()
        p.run()
        p.sync()
        ref = (p.record(), [p.field(w) for w in range(9)], p.step_maxima(), p.snapshots())
    with qp3d.Plan(s, snapflag=1, ntsnap=13, use_graph=True) as p:
        for shot in range(2):
            say(f"graph: replayed shot {shot}")
            p.reset()
            for k0, k1 in ((0, 7), (7, 40), (40, 57)):
                p.run(k0, k1)
                p.sync()
            got = (p.record(), [p.field(w) for w in range(9)], p.step_maxima(), p.snapshots())
            assert np.array_equal(got[0], ref[0])
            for a, b in zip(got[1], ref[1]):
                assert np.array_equal(a, b)
            assert np.array_equal(got[2], ref[2])
            assert [k for k, _ in got[3]] == [k for k, _ in ref[3]] == [0, 13, 26, 39, 52]
            for (_, a), (_, b) in zip(got[3], ref[3]):
                assert np.array_equal(a, b)


# Measured on MI355X for this case (41x45x37 interior, pml 12, 200 steps) against the FP64
# oracle: FP32 1.3e-6, FP16 3.0e-3, BF16 2.5e-2 relative L2 on the gather. The bounds below
# are those figures with a 2x margin: a characterisation that catches a change of the storage
# modes' behaviour, not a parity claim. The precision target (1e-3 against FP64) is
# tests/test_gpu_configs.py::test_c3_16bit_storage_meets_1e3_against_fp64, an expected failure.
LP_TOL = {"fp16": 6e-3, "bf16": 5e-2}


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_low_precision_storage_characterisation(storage):
    """Wavefields and PML state stored in 16 bits, FP32 arithmetic (characterisation, not
    parity): the gather error against the FP64 oracle stays at its measured size, the values
    held are exactly representable in the storage type, and the FP16 scale is taken out of
    everything read back (record, state, snapshots, maxima)."""
    s = qp3d.config_setup(hx=41, hy=45, nz0=37, pml=12, nt=200, seed=3, src_depth=3)
    say(f"storage {storage}")
    o64 = O.modeling3d(s, "f64")
    with qp3d.Plan(s, snapflag=1, ntsnap=100) as p:
        r32 = p.forward()
        g32 = p.step_maxima()
        s32 = p.snapshots()
    with qp3d.Plan(s, storage=storage, snapflag=1, ntsnap=100) as p:
        rec = p.forward()
        g = p.step_maxima()
        u = p.field(0)
        snaps = p.snapshots()
    err = rel_l2(rec, o64)
    say(f"{storage} gather rel-L2 vs FP64 {err:.3e}")
    assert rel_l2(r32, o64) <= 1e-5
    assert 1e-5 < err <= LP_TOL[storage]
    assert rel_l2(g, g32) <= 5 * LP_TOL[storage]
    if storage == "bf16":
        assert not np.any(u.view(np.uint32) & 0xffff)
    else:
        # the plan stores the fields times 2^e with the peak of source*wavelet at 2^7
        e = np.floor(np.log2(128.0 / (np.abs(s.source).max() * np.abs(s.wavelet).max())))
        v = u * np.float32(2.0 ** e)
        assert np.array_equal(v.astype(np.float16).astype(np.float32), v)
    assert [k for k, _ in snaps] == [k for k, _ in s32] == [0, 100]
    assert rel_l2(snaps[1][1], s32[1][1]) <= 2 * LP_TOL[storage]


@pytest.mark.parametrize("gmax_mode", [1, 2])
def test_s_field_snapshots_against_the_oracle(gmax_mode):
    """S-field snapshots (snap_s): at each snapshot step k the scalar operator S the update of
    step k used, over the padded grid, against the oracle's S of the same step; exactly 1
    outside [5,n-5)^3 and 1 everywhere at step 0 (u = 0). With gmax_mode 2, G_k comes from
    the update kernels and the S of the marked cells from the fix-up."""
    s = qp3d.config_setup(hx=37, hy=41, nz0=33, pml=10, nt=101, seed=3, src_depth=14)
    say(f"S field snapshots, gmax_mode {gmax_mode}")
    with qp3d.Plan(s, snapflag=1, ntsnap=50, snap_s=1, gmax_mode=gmax_mode) as p:
        p.forward()
        got = dict(p.snapshots_s())
        snaps = p.snapshots()
    assert sorted(got) == [0, 50, 100] and [k for k, _ in snaps] == [0, 50, 100]
    o = O.OracleState(s)
    rim = np.ones((s.nz, s.nx, s.ny), bool)
    rim[5:-5, 5:-5, 5:-5] = False
    for k in range(s.nt):
        o.step(s.wavelet[k])
        if k in got:
            Sg, So = got[k], o.field("S")
            d = np.abs(Sg.astype(np.float64) - So)
            say(f"step {k}: S < 1 at {(So < 1).sum()} cells, |dS| > 1e-4 at {(d > 1e-4).sum()} of {d.size}, "
                f"max |dS| elsewhere {d[d <= 1e-4].max():.2e}")
            assert np.all(Sg[rim] == 1.0)
            assert s_fields_agree(Sg, So), k
            if k == 0:
                assert np.all(Sg == 1.0)
    assert (got[100] < 0.999).sum() > 2000  # the anisotropic layers are reached


def _hip():
    import ctypes as C
    qp3d.load()  # the library's HIP runtime, whose device pointers the plan recognises
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMalloc.restype = hip.hipMemcpy.restype = hip.hipFree.restype = C.c_int
    return hip


class _DevBuf:
    """A device buffer through the HIP runtime (1 = host to device, 2 = device to host)."""

    def __init__(self, hip, nbytes, host=None):
        import ctypes as C
        self.hip, self.nbytes, self.ptr = hip, nbytes, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), nbytes) == 0
        if host is not None:
            assert hip.hipMemcpy(self.ptr, host.ctypes.data, nbytes, 1) == 0

    def f32(self):
        import ctypes as C
        return C.cast(self.ptr, C.POINTER(C.c_float))

    def get(self, shape, dtype=np.float32):
        out = np.empty(shape, dtype)
        assert out.nbytes == self.nbytes
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_device_pointers_match_host_pointers(storage):
    """The plan API takes arrays in device memory as well as host memory: upload from device
    buffers and read the record, a wavefield, the snapshots and the maxima into device buffers,
    bit-identical to the host-memory path (FP16 storage also takes its scale out on the device)."""
    import ctypes as C
    from qp3d.plan import _tables
    hip = _hip()
    s = qp3d.config_setup(hx=31, hy=29, nz0=27, pml=9, nt=60, seed=4, src_depth=3)
    say(f"device pointers {storage}")
    bufs = []
    try:
        with qp3d.Plan(s, storage=storage, snapflag=1, ntsnap=30) as p:
            rec_h, u_h, g_h, sn_h = p.forward(), p.field(0), p.step_maxima(), p.snapshots()
            arrs = [np.ascontiguousarray(a, np.float32).reshape(-1)
                    for a in [s.vp, s.eps, s.dele, s.source, s.wavelet] + _tables(s)]
            dev = [_DevBuf(hip, a.nbytes, a) for a in arrs]
            bufs += dev
            lib = p.lib
            assert lib.qp3d_plan_upload(p.h, *[d.f32() for d in dev]) == 0
            p.reset()
            p.run()
            p.sync()
            out = {"rec": (_DevBuf(hip, rec_h.nbytes), rec_h.shape), "u": (_DevBuf(hip, u_h.nbytes), u_h.shape),
                   "g": (_DevBuf(hip, g_h.nbytes), g_h.shape), "sn": (_DevBuf(hip, sn_h[1][1].nbytes), sn_h[1][1].shape)}
            bufs += [b for b, _ in out.values()]
            assert lib.qp3d_plan_record(p.h, out["rec"][0].f32()) == 0
            assert lib.qp3d_plan_field(p.h, 0, out["u"][0].f32()) == 0
            assert lib.qp3d_plan_step_maxima(p.h, out["g"][0].f32()) == 0
            k = C.c_int()
            assert lib.qp3d_plan_snapshot(p.h, 1, C.byref(k), out["sn"][0].f32()) == 0
            assert k.value == sn_h[1][0]
            got = {name: b.get(shape) for name, (b, shape) in out.items()}
        assert np.array_equal(got["rec"], rec_h)
        assert np.array_equal(got["u"], u_h)
        assert np.array_equal(got["g"], g_h)
        assert np.array_equal(got["sn"], sn_h[1][1])
    finally:
        for b in bufs:
            b.free()


@pytest.mark.parametrize("gmax_mode", [1, 2])
def test_run_on_a_caller_stream(gmax_mode):
    """qp3d_plan_run_on orders the steps with a caller's stream: after the caller's stream has
    drained, the record equals a plain run's (with gmax_mode 2 the marking pass and the fix-up
    go on the caller's stream as well)."""
    import ctypes as C
    hip = _hip()
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamSynchronize.argtypes = hip.hipStreamDestroy.argtypes = [C.c_void_p]
    hip.hipStreamCreate.restype = hip.hipStreamSynchronize.restype = hip.hipStreamDestroy.restype = C.c_int
    s = qp3d.config_setup(hx=25, hy=27, nz0=23, pml=8, nt=40, seed=6, src_depth=2)
    st = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(st)) == 0
    try:
        with qp3d.Plan(s, gmax_mode=gmax_mode) as p:
            ref = p.forward()
            p.reset()
            assert p.lib.qp3d_plan_run_on(p.h, 0, s.nt, st) == 0
            assert hip.hipStreamSynchronize(st) == 0
            p.sync()
            assert np.array_equal(p.record(), ref)
    finally:
        hip.hipStreamDestroy(st)


# the newest PML slots hold w13/w33/p3 after the rotation, i.e. the oracle's w12/w32/p2
PML_PAIRS = [(2, "w12"), (3, "w11"), (4, "w32"), (5, "w31"), (6, "p2"), (7, "p1"), (8, "w21")]


def test_predicted_threshold_against_the_oracle():
    """gmax_mode 2: the updates take the threshold predicted from the previous steps' G and
    reduce G_k themselves, and the fix-up recomputes the cells whose decision changed. On this
    anisotropic case the prediction without its fix-up misses the 1e-5 bar by 19x
    (tests/test_oracle.py::test_uncorrected_threshold_prediction_misses_the_parity_bar); here
    gathers, wavefields and G per step meet it, over steps on both paths (fix-up from the
    updates' marks, and the marking pass). The PML state is held to gmax_mode 1's own distance
    from the oracle (on this case the w3 recursion of both modes is 3.2e-5 from the oracle's),
    and to 1e-5 of gmax_mode 1; gathers and wavefields agree with gmax_mode 1 to 5e-6."""
    s = qp3d.config_setup(hx=25, hy=27, nz0=23, pml=8, nt=100, seed=1, src_depth=8)
    gpred = []
    rec, st, gm, orec, o, og = run_both(s, predictions=gpred, gmax_mode=2)
    valid = steps_in_band(gm, gpred[0])
    say(f"predicted threshold: {valid.sum()} of {s.nt} steps inside the marked band")
    assert 5 <= valid.sum() <= s.nt - 20
    # the device's predictions (AR fit after 38 steps, ratio rule before) against the host
    # restatement on the same maxima (not bit for bit: log and exp round differently)
    host = np.array([predict_G(gm, k) for k in range(s.nt)], np.float32)
    np.testing.assert_allclose(gpred[0], host, rtol=1e-5, atol=0)
    assert rel_l2(rec, orec) <= TOL
    assert rel_l2(st[0], o.field("u2")) <= TOL
    assert rel_l2(st[1], o.field("u1")) <= TOL
    np.testing.assert_allclose(gm, og, rtol=1e-5, atol=0)
    with qp3d.Plan(s, gmax_mode=1) as p:
        p.reset()
        p.run()
        p.sync()
        rec1, st1 = p.record(), [p.field(w) for w in range(9)]
        with pytest.raises(qp3d.Qp3dError) as ei:
            p.step_predictions()
        assert ei.value.code == -6
    for which, arr in PML_PAIRS:
        ref = _unified(o, s, arr)
        e1, e2, d = rel_l2(st1[which], ref), rel_l2(st[which], ref), rel_l2(st[which], st1[which])
        say(f"{arr}: oracle distance gmax_mode 1 {e1:.2e}, 2 {e2:.2e}; 2 vs 1 {d:.2e}")
        assert e2 <= max(TOL, 1.05 * e1) and d <= 1e-5, (which, arr)
    d = (rel_l2(rec, rec1), rel_l2(st[0], st1[0]), rel_l2(st[1], st1[1]))
    say(f"gmax_mode 2 vs 1: gather {d[0]:.2e}, wavefields {d[1]:.2e} {d[2]:.2e}")
    assert max(d) <= 5e-6


@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
def test_predicted_threshold_graph_replay_and_storage(storage):
    """gmax_mode 2 replayed from the captured graph (runs split at odd steps, snapshot
    steps) is bit-identical to its direct launches, in FP32 and in 16-bit storage; in 16 bits
    it stays within the storage's rounding of gmax_mode 1."""
    s = qp3d.config_setup(hx=29, hy=23, nz0=21, pml=8, nt=64, seed=6, src_depth=6)
    say(f"predicted threshold, graph, {storage}")
    res = {}
    for mode, graph in ((2, False), (2, True), (1, False)):
        with qp3d.Plan(s, storage=storage, gmax_mode=mode, use_graph=graph, snapflag=1, ntsnap=20) as p:
            p.reset()
            for k0, k1 in ((0, 9), (9, 64)):
                p.run(k0, k1)
                p.sync()
            res[mode, graph] = (p.record(), p.field(0), p.step_maxima(), p.snapshots())
    a, b = res[2, False], res[2, True]
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    for (ka, sa), (kb, sb) in zip(a[3], b[3]):
        assert ka == kb and np.array_equal(sa, sb)
    d = rel_l2(a[0], res[1, False][0])
    say(f"gmax_mode 2 vs 1 ({storage}): gather {d:.2e}")
    assert d <= {"fp32": 1e-6, "fp16": 2e-3, "bf16": 2e-2}[storage]
