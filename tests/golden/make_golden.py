"""Writes the golden fixtures of tests/golden/ from the reference itself.

Run in a container that has the reference checkout (it builds oracle/_ref from the
reference's own CPU_function.c, see oracle/Makefile):

    python tests/golden/make_golden.py

Fixtures (data only: inputs and the reference's outputs):
  ref_setup.npz       FD tables, PML profiles, Ricker wavelets and a spatial source from the
                      reference's fdcoff1/fdcoff2/pmlcoff/Source (CPU_function.c:37-169).
  ref_onestep.npz     a random state on a 21x19x17 grid (pml 6) and the reference kernels'
                      grad/absMaxval/scalar_operator/wavefield_update/addsource outputs after
                      one step (CPU_function.c:301-480).
  ref_trajectory.npz  60 steps of the reference's kernels from rest on a small layered VTI
                      shot: the receiver gather, the per-step maxima and the final wavefield.
  ref_output_step0.npz  step 0 of the reference's shipped output/Record_fullwave.dat
                      (run 1, [nt=500][24][24]).
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "gpu-c_3dqmodeling_amd"))

from oracle import oracle as O  # noqa: E402

REFDIR = "/root/reference"


def fp(a):
    return O._fp(a)


def setup_fixture(R):
    c = np.zeros(36, np.float32)
    c2 = np.zeros(36, np.float32)
    R.ref_fdcoff1(fp(c))
    R.ref_fdcoff2(fp(c2))
    out = {"c": c, "c2": c2}
    cases = [(10, 3605.0, 10.0, 10.0, 10.0), (20, 4452.4, 12.5, 12.5, 12.5), (32, 4452.4, 12.5, 10.0, 7.5),
             (7, 2500.0, 5.0, 6.0, 7.0)]
    for n, (pml, vpmax, dx, dy, dz) in enumerate(cases):
        tabs = [np.zeros(pml, np.float32) for _ in range(6)]
        R.ref_pmlcoff(pml, vpmax, dx, dy, dz, *[fp(t) for t in tabs])
        out[f"pml{n}_args"] = np.array([pml, vpmax, dx, dy, dz], np.float64)
        for name, t in zip(["dlr", "ddlr", "dtb", "ddtb", "dfb", "ddfb"], tabs):
            out[f"pml{n}_{name}"] = t
    # Ricker wavelets and a spatial source on a small padded grid
    for n, (fm, dt, nt) in enumerate([(25.0, 0.001, 300), (30.0, 0.001, 500), (20.0, 0.0005, 400)]):
        wav = np.zeros(nt, np.float32)
        t0 = float(np.float32(1.0 / fm))
        R.ref_source(0, 0, 0, fm, 1.0, 1.0, dt, 10.0, 10.0, 10.0, t0, nt, 1, 1, 1,
                     fp(np.zeros(1, np.float32)), fp(wav))
        out[f"wav{n}_args"] = np.array([fm, dt, nt, t0], np.float64)
        out[f"wav{n}"] = wav
    nx, ny, nz = 13, 11, 9
    src = np.zeros(nx * ny * nz, np.float32)
    wav = np.zeros(10, np.float32)
    args = (6, 5, 4, 30.0, 1.5, 0.02, 0.001, 10.0, 12.0, 8.0, float(np.float32(1 / 30.0)), 10, nx, ny, nz)
    R.ref_source(*args, fp(src), fp(wav))
    out["src_args"] = np.array(args, np.float64)
    out["src"] = src
    # extend_model on a random interior
    rng = np.random.default_rng(7)
    pml, hx, hy, hz = 4, 5, 6, 3
    inner = rng.standard_normal(hx * hy * hz).astype(np.float32)
    ext = np.zeros((hx + 2 * pml) * (hy + 2 * pml) * (hz + 2 * pml), np.float32)
    R.ref_extend_model(fp(inner), fp(ext), hx + 2 * pml, hy + 2 * pml, hz + 2 * pml, pml)
    out["ext_in"] = inner
    out["ext_out"] = ext
    out["ext_dims"] = np.array([hx, hy, hz, pml])
    np.savez_compressed(os.path.join(HERE, "ref_setup.npz"), **out)


def onestep_fixture():
    import qp3d
    rng = np.random.default_rng(11)
    hx, hy, nz0, pml = 9, 7, 5, 6
    vp1 = (2000 + 1500 * rng.random((nz0, hx, hy))).astype(np.float32)
    ep1 = (0.25 * rng.random((nz0, hx, hy))).astype(np.float32)
    de1 = np.minimum(ep1, (0.15 * rng.random((nz0, hx, hy))).astype(np.float32))
    s = qp3d.build_setup(vp1, ep1, de1, pml, 4, 12.5, 12.5, 12.5, 0.001, 25.0, src_depth=1)
    r = O.RefState(s)
    n = s.nx * s.ny * s.nz
    state = {"u1": rng.standard_normal(n).astype(np.float32) * 1e-2,
             "u2": rng.standard_normal(n).astype(np.float32) * 1e-2}
    for k, v in state.items():
        r.set_field(k, v.reshape(s.nz, s.nx, s.ny))
    aux_in = {}
    for face in O.FACES:
        for arr in O.AUX:
            a = (rng.standard_normal(r._aux_len(O.FACES.index(face))) * 1e-3).astype(np.float32)
            r.set_aux(face, arr, a)
            aux_in[f"aux_{face}_{arr}"] = a
    w = float(s.wavelet[2])
    r.step(w)
    out = {"model_vp": vp1, "model_eps": ep1, "model_del": de1, "w": np.float32(w), "G": np.float64(r.G)}
    out.update({f"in_{k}": v for k, v in state.items()})
    out.update({f"in_{k}": v for k, v in aux_in.items()})
    for f in ("u1", "u2", "ux", "uy", "uz", "S"):
        out[f"out_{f}"] = r.field(f).copy()
    for face in O.FACES:
        for arr in ("w11", "w12", "w21", "w31", "w32", "p1", "p2"):
            out[f"out_aux_{face}_{arr}"] = r.aux(face, arr).copy()
    np.savez_compressed(os.path.join(HERE, "ref_onestep.npz"), **out)


def trajectory_fixture():
    import qp3d
    s = qp3d.config_setup(hx=13, hy=11, nz0=12, pml=8, nt=60, seed=5, src_depth=2)
    r = O.RefState(s)
    rec = np.zeros((s.nt, s.hx, s.hy), np.float32)
    G = np.zeros(s.nt, np.float64)
    for k in range(s.nt):
        r.step(s.wavelet[k])
        G[k] = r.G
        rec[k] = r.field("u2")[s.sz, s.pml:s.nx - s.pml, s.pml:s.ny - s.pml]
    np.savez_compressed(os.path.join(HERE, "ref_trajectory.npz"), record=rec, G=G, u2=r.field("u2").copy(),
                        args=np.array([13, 11, 12, 8, 60, 5, 2]))


def output_fixture():
    a = np.fromfile(os.path.join(REFDIR, "output", "Record_fullwave.dat"), dtype="<f4")
    step0 = a[:24 * 24].reshape(24, 24)
    np.savez_compressed(os.path.join(HERE, "ref_output_step0.npz"), step0=step0,
                        meta=np.array([500, 24, 24, 10, 30, 10.0]))


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    R = O.ref_lib()
    setup_fixture(R)
    onestep_fixture()
    trajectory_fixture()
    output_fixture()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
