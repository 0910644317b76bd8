"""Pins the CPU oracle: bit-exact against the reference's own CPU kernels (oracle/_ref, when
built here), against golden vectors written from them (tests/golden), against the
reference's shipped output file, and against closed forms and invariants that a dropped
term, sign or index would break."""
import dataclasses
import os

import numpy as np
import pytest

from conftest import predict_G, rel_l2, steps_in_band

import qp3d
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref needs the reference checkout")


def bits(a):
    return np.asarray(a).view(np.uint32)


def small_setup(hx=13, hy=11, nz0=9, pml=7, nt=30, seed=0, depth=1):
    return qp3d.config_setup(hx=hx, hy=hy, nz0=nz0, pml=pml, nt=nt, seed=seed, src_depth=depth)


def compare_all(a, b):
    for f in O.FIELDS:
        assert np.array_equal(bits(a.field(f)), bits(b.field(f))), f
    for face in O.FACES:
        for arr in O.AUX:
            assert np.array_equal(bits(a.aux(face, arr)), bits(b.aux(face, arr))), (face, arr)


@needs_ref
def test_bitexact_vs_reference_kernels_from_rest():
    s = small_setup(hx=17, hy=13, nz0=11, pml=8, nt=45, seed=2, depth=2)
    o, r = O.OracleState(s), O.RefState(s)
    for k in range(s.nt):
        o.step(s.wavelet[k])
        r.step(s.wavelet[k])
        assert o.G == r.G
    compare_all(o, r)
    assert np.abs(o.field("u2")).max() > 0


@needs_ref
def test_bitexact_vs_reference_kernels_random_state():
    """Random wavefields and random PML state in every face array, a few steps."""
    rng = np.random.default_rng(5)
    s = small_setup(hx=12, hy=14, nz0=10, pml=9, nt=6, seed=4, depth=3)
    o, r = O.OracleState(s), O.RefState(s)
    n = s.nx * s.ny * s.nz
    for f in ("u1", "u2", "u3"):
        a = (rng.standard_normal(n) * 1e-2).astype(np.float32).reshape(s.nz, s.nx, s.ny)
        o.set_field(f, a)
        r.set_field(f, a)
    for face in O.FACES:
        for arr in O.AUX:
            a = (rng.standard_normal(o._aux_len(O.FACES.index(face))) * 1e-3).astype(np.float32)
            o.set_aux(face, arr, a)
            r.set_aux(face, arr, a)
    for k in range(s.nt):
        o.step(s.wavelet[k])
        r.step(s.wavelet[k])
    compare_all(o, r)


def test_onestep_golden_vectors():
    g = np.load(os.path.join(GOLD, "ref_onestep.npz"))
    s = qp3d.build_setup(g["model_vp"], g["model_eps"], g["model_del"], 6, 4, 12.5, 12.5, 12.5, 0.001, 25.0,
                         src_depth=1)
    o = O.OracleState(s)
    shp = (s.nz, s.nx, s.ny)
    o.set_field("u1", g["in_u1"].reshape(shp))
    o.set_field("u2", g["in_u2"].reshape(shp))
    for face in O.FACES:
        for arr in O.AUX:
            o.set_aux(face, arr, g[f"in_aux_{face}_{arr}"])
    o.step(float(g["w"]))
    assert o.G == float(g["G"])
    for f in ("u1", "u2", "ux", "uy", "uz", "S"):
        assert np.array_equal(bits(o.field(f)), bits(g[f"out_{f}"])), f
    for face in O.FACES:
        for arr in ("w11", "w12", "w21", "w31", "w32", "p1", "p2"):
            assert np.array_equal(bits(o.aux(face, arr)), bits(g[f"out_aux_{face}_{arr}"])), (face, arr)


def test_trajectory_golden_vectors():
    g = np.load(os.path.join(GOLD, "ref_trajectory.npz"))
    hx, hy, nz0, pml, nt, seed, depth = [int(v) for v in g["args"]]
    s = qp3d.config_setup(hx=hx, hy=hy, nz0=nz0, pml=pml, nt=nt, seed=seed, src_depth=depth)
    o = O.OracleState(s)
    for k in range(nt):
        o.step(s.wavelet[k])
        assert o.G == g["G"][k]
        rec = o.field("u2")[s.sz, pml:s.nx - pml, pml:s.ny - pml]
        assert np.array_equal(bits(rec), bits(g["record"][k])), k
    assert np.array_equal(bits(o.field("u2")), bits(g["u2"]))


def test_modeling3d_loop_matches_stepper():
    s = small_setup(nt=25)
    rec = O.modeling3d(s, "f32", nthreads=2)
    o = O.OracleState(s, nthreads=3)
    for k in range(s.nt):
        o.step(s.wavelet[k])
        assert np.array_equal(bits(rec[k]), bits(o.field("u2")[s.sz, s.pml:-s.pml, s.pml:-s.pml]))


def test_reference_output_step0_fixture():
    """Record_fullwave.dat of the reference (24x24 window, pml 10, dx 10 m, fm 30 Hz): after
    step 0 only the source cell (11, 11) is non-zero and equals amp * w[0]."""
    g = np.load(os.path.join(GOLD, "ref_output_step0.npz"))["step0"]
    rng = np.random.default_rng(0)
    vp = (2000 + 1000 * rng.random((24, 24, 24))).astype(np.float32)
    ep = (0.2 * rng.random((24, 24, 24))).astype(np.float32)
    de = (0.5 * ep).astype(np.float32)
    s = qp3d.build_setup(vp, ep, de, 10, 3, 10.0, 10.0, 10.0, 0.001, 30.0)
    rec = O.modeling3d(s, "f32", nt=1)
    # equal as values: the file holds -0.0 at the four neighbours of the source, the sign an
    # FMA-contracted `u3 += source*wavelet` gives to the underflowed denormal product (the
    # CUDA build); the CPU path's separate multiply and add gives +0.0 there.
    assert np.array_equal(rec[0], g)
    assert np.count_nonzero(g) == 1 and g[11, 11] == np.float32(-9.6925104e-4)
    assert np.array_equal(np.argwhere(np.signbit(g)), [[10, 11], [11, 10], [11, 11], [11, 12], [12, 11]])


# ---- closed forms ---------------------------------------------------------------------

def test_fd_coefficient_moments():
    """With the rounded decimal coefficients, d/dx of a ramp is 2*sum(m c1_m) = 0.9963 and
    d2/dx2 of x^2 is 2*sum(m^2 c2_m) = 2*0.9991 (exact expected values, not 1)."""
    c, c2 = qp3d.fd_tables()
    m = np.arange(1, 6)
    assert abs(2 * np.sum(m * c[31:36].astype(np.float64)) - 0.9963) < 1e-6
    assert abs(np.sum(m * m * c2[31:36].astype(np.float64)) - 0.9991) < 1e-6


def _uniform_setup(nx=31, ny=29, nz=27, pml=8, vp=2000.0, eps=0.0, dele=0.0, nt=3):
    hx, hy, hz = nx - 2 * pml, ny - 2 * pml, nz - 2 * pml
    a = lambda v: np.full((hz, hx, hy), v, np.float32)  # noqa: E731
    return qp3d.build_setup(a(vp), a(eps), a(dele), pml, nt, 10.0, 12.0, 8.0, 0.001, 25.0)


def test_grad_of_linear_ramps():
    s = _uniform_setup()
    o = O.OracleState(s, "f64")
    z, x, y = np.meshgrid(np.arange(s.nz), np.arange(s.nx), np.arange(s.ny), indexing="ij")
    u = (3.0 * x * s.dx - 2.0 * y * s.dy + 0.5 * z * s.dz)
    o.set_field("u2", u)
    o.step(0.0)
    ux, uy, uz = o.field("ux"), o.field("uy"), o.field("uz")
    inner = (slice(5, -5),) * 3
    np.testing.assert_allclose(ux[inner], 3.0 * 0.9963, rtol=1e-6)
    np.testing.assert_allclose(uy[inner], -2.0 * 0.9963, rtol=1e-6)
    np.testing.assert_allclose(uz[inner], 0.5 * 0.9963, rtol=1e-6)
    assert not np.any(ux[:5]) and not np.any(uy[:, :, -5:])


def test_interior_update_of_quadratic():
    """u1 = u2 = x^2 (metres): the interior update adds dt^2 vp^2 * 2*0.9991 per step."""
    s = _uniform_setup(vp=2000.0)
    o = O.OracleState(s, "f64")
    z, x, y = np.meshgrid(np.arange(s.nz), np.arange(s.nx), np.arange(s.ny), indexing="ij")
    u = (x * s.dx) ** 2 * 1e-3
    o.set_field("u1", u)
    o.set_field("u2", u)
    o.step(0.0)
    p = s.pml
    d = o.field("u2")[p:-p, p:-p, p:-p] - u[p:-p, p:-p, p:-p]
    expect = s.dt ** 2 * np.float64(np.float32(2000.0)) ** 2 * 2 * 0.9991 * 1e-3
    np.testing.assert_allclose(d, expect, rtol=1e-6)


def test_ricker_first_sample():
    """t0 = 1/fm makes a = pi^2 at k = 0 for every fm: w[0] = (1-2 pi^2) e^(-pi^2)."""
    for fm in (20.0, 25.0, 30.0):
        s = _uniform_setup(nt=5)
        s2 = qp3d.build_setup(s.vp[8:-8, 8:-8, 8:-8], s.eps[8:-8, 8:-8, 8:-8], s.dele[8:-8, 8:-8, 8:-8], 8, 5,
                              10.0, 10.0, 10.0, 0.001, fm)
        assert abs(s2.wavelet[0] - (1 - 2 * np.pi ** 2) * np.exp(-np.pi ** 2)) < 2e-9


def test_pml_profile_formula():
    s = _uniform_setup(vp=2500.0, eps=0.1)
    L = s.pml * s.dx
    w = (np.arange(s.pml) + 1) * s.dx
    vpmax = (1 + 2 * 0.1) * 2500.0
    d = 4 * vpmax / (2 * L) * (w / L) ** 3 * np.log(1e5)
    dd = 6 * vpmax * w * w / L ** 4 * np.log(1e5)
    np.testing.assert_allclose(s.dlr, d, rtol=2e-6)
    np.testing.assert_allclose(s.ddlr, dd, rtol=2e-6)
    Lz = s.pml * s.dz
    np.testing.assert_allclose(s.dtb, 4 * vpmax / (2 * Lz) * (((np.arange(s.pml) + 1) * s.dz) / Lz) ** 3 *
                               np.log(1e5), rtol=2e-6)


# ---- invariants -----------------------------------------------------------------------

def test_eps_equal_delta_gives_unit_S():
    s = _uniform_setup(vp=3000.0, eps=0.2, dele=0.2, nt=12)
    o = O.OracleState(s)
    for k in range(s.nt):
        o.step(s.wavelet[k])
    assert np.abs(o.field("u2")).max() > 0
    assert np.all(o.field("S") == 1.0)


def test_S_below_one_in_anisotropic_medium():
    s = _uniform_setup(vp=3000.0, eps=0.2, dele=0.05, nt=12)
    o = O.OracleState(s)
    for k in range(s.nt):
        o.step(s.wavelet[k])
    S = o.field("S")
    assert S.min() < 0.99 and S.max() == 1.0


def _acoustic_numpy(s, nsteps):
    """Textbook 10th-order acoustic leapfrog u3 = 2u2 - u1 + dt^2 vp^2 lap(u) + src w, computed
    independently in float64 on the interior."""
    c2 = s.c2[31:36].astype(np.float64)
    u1 = np.zeros((s.nz, s.nx, s.ny))
    u2 = np.zeros_like(u1)
    vp = s.vp.astype(np.float64)
    src = s.source.astype(np.float64)
    for k in range(nsteps):
        lap = np.zeros_like(u2)
        c = u2[5:-5, 5:-5, 5:-5]
        for m in range(1, 6):
            lap[5:-5, 5:-5, 5:-5] += c2[m - 1] * (
                (u2[5:-5, 5 + m:u2.shape[1] - 5 + m, 5:-5] - 2 * c + u2[5:-5, 5 - m:u2.shape[1] - 5 - m, 5:-5])
                / np.float64(np.float32(s.dx) * np.float32(s.dx)) +
                (u2[5:-5, 5:-5, 5 + m:u2.shape[2] - 5 + m] - 2 * c + u2[5:-5, 5:-5, 5 - m:u2.shape[2] - 5 - m])
                / np.float64(np.float32(s.dy) * np.float32(s.dy)) +
                (u2[5 + m:u2.shape[0] - 5 + m, 5:-5, 5:-5] - 2 * c + u2[5 - m:u2.shape[0] - 5 - m, 5:-5, 5:-5])
                / np.float64(np.float32(s.dz) * np.float32(s.dz)))
        dt = np.float64(np.float32(s.dt))
        u3 = 2 * u2 - u1 + dt ** 2 * vp ** 2 * lap + src * np.float64(s.wavelet[k])
        u1, u2 = u2, u3
    return u2


def test_isotropic_reduces_to_acoustic_scheme():
    """eps = delta = 0: the qP operator is the plain acoustic one while the wavefield is still
    inside the interior (S is 1 and no PML cell has been reached)."""
    s = _uniform_setup(nx=81, ny=81, nz=81, pml=8, vp=2000.0, nt=6)
    src = np.zeros_like(s.source)
    src[40, 40, 40] = 1.0  # the stencil reach after 6 steps (30 cells) stays off the PML
    s.source = src
    o = O.OracleState(s, "f64")
    for k in range(s.nt):
        o.step(s.wavelet[k])
    ref = _acoustic_numpy(s, s.nt)
    got = o.field("u2")
    assert np.abs(got).max() > 0
    p = s.pml
    assert not np.any(got[:p]) and not np.any(got[:, :p]) and not np.any(got[:, :, :p])
    assert rel_l2(got, ref) < 1e-12


def test_rim_stays_zero_and_step0_is_impulse():
    s = small_setup(nt=10, depth=2)
    o = O.OracleState(s)
    o.step(s.wavelet[0])
    u = o.field("u2")
    sx, sy, sz = s.src_cell
    nz_ = np.argwhere(u != 0)
    assert nz_.tolist() == [[sz, sx, sy]]
    assert u[sz, sx, sy] == np.float32(s.wavelet[0])
    for k in range(1, s.nt):
        o.step(s.wavelet[k])
    u = o.field("u2")
    for sl in (u[:5], u[-5:], u[:, :5], u[:, -5:], u[:, :, :5], u[:, :, -5:]):
        assert not np.any(sl)


def test_f64_and_f32_oracles_agree():
    s = small_setup(hx=15, hy=15, nz0=13, pml=8, nt=60, seed=3)
    a = O.modeling3d(s, "f32")
    b = O.modeling3d(s, "f64")
    assert rel_l2(a, b) < 1e-5


def _export(o, s):
    """Full-grid state of an oracle (the layout the HIP plan uses)."""
    from test_gpu_parity import _unified
    u_new, u_old = o.field("u2").copy(), o.field("u1").copy()
    aux = [_unified(o, s, a).astype(np.float32) for a in ("w12", "w11", "w32", "w31", "p2", "p1", "w21")]
    return [u_new, u_old] + aux


def test_import_state_round_trip():
    s = small_setup(hx=14, hy=12, nz0=10, pml=8, nt=40, seed=6, depth=1)
    a = O.OracleState(s)
    for k in range(25):
        a.step(s.wavelet[k])
    b = O.OracleState(s)
    b.import_state(*_export(a, s))
    for k in range(25, 40):
        a.step(s.wavelet[k])
        b.step(s.wavelet[k])
    for f in ("u1", "u2"):
        assert np.array_equal(bits(a.field(f)), bits(b.field(f)))


def test_uncorrected_threshold_prediction_misses_the_parity_bar():
    """Sensitivity of the GPU parity test of gmax_mode 2 (predicted threshold plus fix-up):
    the oracle with a planted mistake, S thresholded at the predicted G instead of G_k (the
    prediction without its fix-up), moves the gather of this anisotropic case far beyond the
    1e-5 bar, and the case holds steps on both sides of the band test (fix-up from the update's
    marks, and the marking pass)."""
    s = qp3d.config_setup(hx=25, hy=27, nz0=23, pml=8, nt=100, seed=1, src_depth=8)
    a, b = O.OracleState(s), O.OracleState(s)
    ga, ra, rb = [], [], []
    for k in range(s.nt):
        a.step(s.wavelet[k])
        ga.append(np.float32(a.G))
        b.step_planted_G(s.wavelet[k], float(predict_G(ga, k)))
        ra.append(a.field("u2")[s.sz].copy())
        rb.append(b.field("u2")[s.sz].copy())
    ga = np.array(ga, np.float32)
    valid = steps_in_band(ga, [predict_G(ga, k) for k in range(s.nt)])
    assert 5 <= valid.sum() <= s.nt - 20
    assert rel_l2(np.array(rb), np.array(ra)) > 10 * 1e-5  # measured 1.9e-4


def test_frozen_decisions_record_and_replay():
    """The decision hook of the oracle: recording a run's per-cell threshold decisions changes
    nothing, replaying them reproduces the run bit for bit (FP32 and FP64), and replaying other
    decisions (every cell below the threshold, S = 1) moves the record, so the hook is what sets
    S's branch."""
    s = qp3d.config_setup(hx=21, hy=23, nz0=19, pml=8, nt=60, seed=0, src_depth=3)
    for prec in ("f32", "f64"):
        plain = O.modeling3d(s, prec)
        o = O.OracleState(s, prec)
        bits = O.decision_bits(s)
        o.set_decisions(bits, 1)
        rec = o.run_record()
        o.close()
        assert np.array_equal(rec, plain)
        assert 1000 < np.unpackbits(bits).sum() < bits.size * 8
        assert np.array_equal(O.record_frozen(s, bits, prec), plain)
        assert rel_l2(O.record_frozen(s, np.zeros_like(bits), prec), plain) > 1e-4


def test_frozen_decision_misfit_is_smooth():
    """With the decisions frozen the misfit is a smooth function of the model, so central
    differences converge quadratically in the step: on an anisotropic model the estimates at h,
    h/2 and h/4 differ by a ratio close to 4; without freezing, threshold crossings inside a
    step make the differences jump instead (measured: ratios 3.8 frozen, -0.02 not)."""
    s = qp3d.config_setup(hx=26, hy=24, nz0=22, pml=8, nt=120, seed=4, src_depth=3)
    obs = O.modeling3d(s, "f64") * 0.9
    o = O.OracleState(s, "f64")
    bits = O.decision_bits(s)
    o.set_decisions(bits, 1)
    o.run_record()
    o.close()
    rng = np.random.default_rng(1)
    from scipy.ndimage import gaussian_filter
    d = gaussian_filter(rng.standard_normal(s.vp.shape), 2.0)
    d /= np.abs(d).max()

    def fd(h, frozen):
        mp = (s.eps.astype(np.float64) + h * d).astype(np.float32)
        mm = (s.eps.astype(np.float64) - h * d).astype(np.float32)
        run = (lambda m: O.record_frozen(m, bits)) if frozen else (lambda m: O.modeling3d(m, "f64"))
        jp = 0.5 * np.sum((run(dataclasses.replace(s, eps=mp)) - obs) ** 2)
        jm = 0.5 * np.sum((run(dataclasses.replace(s, eps=mm)) - obs) ** 2)
        # per unit of the perturbation actually applied (float32 rounding of the models)
        return (jp - jm) / np.sum((mp.astype(np.float64) - mm.astype(np.float64)) * d)

    def ratio(frozen):
        d1, d2, d4 = (fd(h, frozen) for h in (4e-3, 2e-3, 1e-3))
        return (d1 - d2) / (d2 - d4)

    assert 3.0 < ratio(True) < 5.0
    assert abs(ratio(False) - 4.0) > 1.0  # the threshold crossings of this case break it
