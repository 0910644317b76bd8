"""The drop-in driver qp3d_main (Main.c:24-246 on top of the library) against the same shot
loop run with the CPU oracle: parameter file, model windows read from model files, the direct
wave, full-wave and muted records appended per shot, the xz / yz slices of the first shot and
the snapshots, compared file by file (same sizes; FP32 values within relative L2 1e-5)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import rel_l2, s_fields_agree

import qp3d
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5
MAIN = os.path.join(os.path.dirname(qp3d.LIB_PATH), "qp3d_main")

PAR = """Time sampling rate[dt-ms]--record time length[nt]
0.001\t{nt}\t\t{ntsnap}
X-direction distance--y-direction distance--z-direction distance[zmax-m]
{Nx0}\t  {Ny0}     {Nz0}
X-direction sampling internal--y-direction sampling internal--z-direction sampling internal[dz-m]
12.5\t12.5   \t12.5
Ricker wavelet dominant frequency--source amplitude--source decay coefficient[alp]
25.0\t1    \t1
First source position at x-direction[spx-m] y-direction[spy-m]and z-direction[spz-m]
{sx0}\t\t{sy0}     \t0
First geophone position at x-direction[gx-m] y-direction[gy-m]and z-direction[gz-m]
0  \t\t0  \t\t0
Nshot(x,y)--shot_interval(x,y)
{nxs}  \t\t{nys}  \t\t{dxs}  \t{dys}
Offset[m]-the distance of offset in x and y
{hx}  \t{hy}
The number of PML absorbing boundary layers[npml]
{pml}
Model-parameters-vp-epsilon-delta-filename
model/vp.dat\tmodel/epsilon.dat\tmodel/delta.dat
Record-fullwave-directwave-mute-filename
output/Record_fullwave.dat   \toutput/Record_direct.dat   output/Record_mute.dat
Record-fullwave-xoz and yoz-slice
output/recordxz.dat   \toutput/recordyz.dat
Snapshot-fullwve-directwave-filename
output/snapshot_direct  \toutput/snapshot_fullwave
Single-shot or multi-shots (single = 0, multi = 1)
1
Mute directwave or not(yes = 1, no = 0)
1
Save fullwave wavefield snapshot or not(yes = 1, no = 0)
1
Save directwave wavefield snapshot or not(yes = 1, no = 0)
0
"""


def oracle_run(d, p, model):
    """The shot loop of Main.c:163-247 with the oracle's modeling3d (homogeneous 2500 m/s
    direct-wave model, as the driver documents)."""
    os.makedirs(os.path.join(d, "output"))
    nx, ny, nz = p["hx"] + 2 * p["pml"], p["hy"] + 2 * p["pml"], p["Nz0"] + 2 * p["pml"]
    z = np.zeros((p["Nz0"], p["hx"], p["hy"]), np.float32)
    sd = qp3d.build_setup(z + 2500.0, z, z, p["pml"], p["nt"], 12.5, 12.5, 12.5, 0.001, 25.0)
    rec_dir = O.modeling3d(sd, "f32")
    rec_dir.tofile(os.path.join(d, "output/Record_direct.dat"))
    for ix in range(p["nxs"]):
        for iy in range(p["nys"]):
            sx, sy = p["sx0"] + ix * p["dxs"], p["sy0"] + iy * p["dys"]
            fnx, fny = sx - (p["hx"] - 1) // 2, sy - (p["hy"] - 1) // 2
            win = [qp3d.file_to_grid(np.ascontiguousarray(m[fny:fny + p["hy"], fnx:fnx + p["hx"], :])) for m in model]
            s = qp3d.build_setup(*win, p["pml"], p["nt"], 12.5, 12.5, 12.5, 0.001, 25.0)
            full = O.modeling3d(s, "f32", snapflag=1, ntsnap=p["ntsnap"],
                                snap_file=os.path.join(d, "output/snapshot_fullwave"))
            with open(os.path.join(d, "output/Record_fullwave.dat"), "ab") as f:
                full.tofile(f)
            with open(os.path.join(d, "output/Record_mute.dat"), "ab") as f:
                (full - rec_dir).tofile(f)
            if ix == 0 and iy == 0:
                sx1, sy1 = (p["hx"] - 1) // 2, (p["hy"] - 1) // 2
                np.ascontiguousarray(full[:, :, sy1]).tofile(os.path.join(d, "output/recordxz.dat"))
                np.ascontiguousarray(full[:, sx1, :]).tofile(os.path.join(d, "output/recordyz.dat"))


@pytest.mark.parametrize("gmax_mode", ["", "2"])
def test_qp3d_main_matches_the_oracle_driver(tmp_path, monkeypatch, gmax_mode):
    """Every output file of qp3d_main against the oracle driver's; with QP3D_GMAX_MODE=2 the
    drop-in call takes G from the update kernels (predicted threshold and fix-up)."""
    assert os.path.exists(MAIN), "qp3d_main is not built"
    if gmax_mode:
        monkeypatch.setenv("QP3D_GMAX_MODE", gmax_mode)
    p = dict(nt=80, ntsnap=40, Nx0=40, Ny0=36, Nz0=24, sx0=14, sy0=15, nxs=2, nys=1, dxs=8, dys=6, hx=21, hy=19,
             pml=10)
    model = qp3d.genmodel(p["Nx0"], p["Ny0"], p["Nz0"], 5)  # file layout [y][x][z]
    g, c = tmp_path / "gpu", tmp_path / "cpu"
    os.makedirs(g / "model")
    os.makedirs(g / "output")
    for name, a in zip(("vp", "epsilon", "delta"), model):
        a.tofile(g / "model" / f"{name}.dat")
    (g / "parameter.txt").write_text(PAR.format(**p))
    print("running qp3d_main", flush=True)
    r = subprocess.run([MAIN, "parameter.txt"], cwd=g, capture_output=True, text=True, timeout=120)
    print(r.stdout[-500:], r.stderr[-500:], flush=True)
    assert r.returncode == 0
    print("running the oracle driver", flush=True)
    oracle_run(str(c), p, model)
    names = sorted(os.listdir(c / "output"))
    assert sorted(os.listdir(g / "output")) == names
    assert "snapshot_fullwave40.dat" in names and "snapshot_fullwave_xz0.dat" in names
    for n in names:
        a = np.fromfile(g / "output" / n, np.float32)
        b = np.fromfile(c / "output" / n, np.float32)
        assert a.size == b.size, n
        if np.any(b):
            assert rel_l2(a, b) <= TOL, n
        else:
            assert not np.any(a), n
    R = p["nt"] * p["hx"] * p["hy"]
    assert os.path.getsize(g / "output/Record_fullwave.dat") == 4 * R * p["nxs"] * p["nys"]


def test_qp3d_main_writes_s_field_snapshots(tmp_path, monkeypatch):
    """QP3D_SNAP_S=1: the drop-in call also writes the S field of every snapshot step over the
    padded grid and its three centre slices, under the names of the reference's commented-out
    writer (3dPureqPModeling_single_shot.cu:381-409); compared with the oracle driver's files."""
    monkeypatch.setenv("QP3D_SNAP_S", "1")
    p = dict(nt=61, ntsnap=30, Nx0=30, Ny0=28, Nz0=24, sx0=14, sy0=13, nxs=1, nys=1, dxs=8, dys=6, hx=21, hy=19,
             pml=10)
    model = qp3d.genmodel(p["Nx0"], p["Ny0"], p["Nz0"], 6)
    g, c = tmp_path / "gpu", tmp_path / "cpu"
    os.makedirs(g / "model")
    os.makedirs(g / "output")
    for name, a in zip(("vp", "epsilon", "delta"), model):
        a.tofile(g / "model" / f"{name}.dat")
    (g / "parameter.txt").write_text(PAR.format(**p))
    print("running qp3d_main with QP3D_SNAP_S=1", flush=True)
    r = subprocess.run([MAIN, "parameter.txt"], cwd=g, capture_output=True, text=True, timeout=120)
    print(r.stdout[-500:], r.stderr[-500:], flush=True)
    assert r.returncode == 0
    oracle_run(str(c), p, model)
    names = sorted(os.listdir(c / "output"))
    assert sorted(os.listdir(g / "output")) == names
    nx, ny, nz = p["hx"] + 2 * p["pml"], p["hy"] + 2 * p["pml"], p["Nz0"] + 2 * p["pml"]
    sizes = {"S": nz * nx * ny, "Sxz": nz * nx, "Syz": nz * ny, "Sxy": nx * ny}
    for k in (0, 30, 60):
        for tag, n in sizes.items():
            name = f"snapshot_fullwave_{tag}{k}.dat"
            assert name in names
            a = np.fromfile(g / "output" / name, np.float32)
            b = np.fromfile(c / "output" / name, np.float32)
            assert a.size == b.size == n, name
            assert s_fields_agree(a, b), name
    for n in names:
        if "_S" not in n:
            a = np.fromfile(g / "output" / n, np.float32)
            b = np.fromfile(c / "output" / n, np.float32)
            assert a.size == b.size and (rel_l2(a, b) <= TOL if np.any(b) else not np.any(a)), n


REF_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")
REF_MAIN_GPU = os.path.join(REF_DIR, "ref_main_qp3d")
REF_MAIN_CPU = os.path.join(REF_DIR, "ref_main_oracle")


@pytest.mark.skipif(not (os.path.exists(REF_MAIN_GPU) and os.path.exists(REF_MAIN_CPU)),
                    reason="the reference's driver is built (make -C oracle ref) only where its checkout is")
def test_reference_driver_linked_unchanged_against_the_library(tmp_path):
    """The reference's own Main.c (Main.c:24-246), compiled unchanged with g++ and linked against
    libqp3d.so through the C++-linkage modeling3d, against the same objects linked against the
    FP32 oracle: the link-time swap the reference makes between its CPU and GPU bodies
    (Makefile:8, compile_GPU.sh:11). Every output file: same names and sizes, values within
    relative L2 1e-5. This driver also runs the reference's direct-wave call with its
    all-zero model (Main.c:163-175), where the oracle and the GPU agree too."""
    p = dict(nt=70, ntsnap=30, Nx0=34, Ny0=30, Nz0=22, sx0=15, sy0=14, nxs=2, nys=1, dxs=6, dys=6, hx=19, hy=17,
             pml=9)
    model = qp3d.genmodel(p["Nx0"], p["Ny0"], p["Nz0"], 11)
    dirs = {}
    for tag, exe in (("gpu", REF_MAIN_GPU), ("cpu", REF_MAIN_CPU)):
        d = tmp_path / tag
        for sub in ("input", "model", "output"):
            os.makedirs(d / sub)
        for name, a in zip(("vp", "epsilon", "delta"), model):
            a.tofile(d / "model" / f"{name}.dat")
        (d / "input" / "parameter.txt").write_text(PAR.format(**p))
        print(f"running the reference driver linked against {'libqp3d.so' if tag == 'gpu' else 'the oracle'}",
              flush=True)
        r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=180)
        print(r.stdout[-300:], r.stderr[-300:], flush=True)
        assert r.returncode == 0, tag
        dirs[tag] = d
    names = sorted(os.listdir(dirs["cpu"] / "output"))
    assert sorted(os.listdir(dirs["gpu"] / "output")) == names
    assert {"Record_direct.dat", "Record_fullwave.dat", "Record_mute.dat", "snapshot_fullwave30.dat"} <= set(names)
    for n in names:
        a = np.fromfile(dirs["gpu"] / "output" / n, np.float32)
        b = np.fromfile(dirs["cpu"] / "output" / n, np.float32)
        assert a.size == b.size and a.size > 0, n
        if np.any(b):
            assert rel_l2(a, b) <= TOL, n
        else:
            assert not np.any(a), n


def test_void_entry_points_through_ctypes():
    """The legacy void modeling3d, with C linkage and with C++ linkage (the mangled symbol a
    g++-built driver binds), called through ctypes against the oracle."""
    import ctypes as C
    from qp3d._lib import fptr
    s = qp3d.config_setup(hx=23, hy=21, nz0=19, pml=9, nt=70, seed=9, src_depth=2)
    orec = O.modeling3d(s, "f32")
    lib = C.CDLL(qp3d.LIB_PATH)
    F = C.POINTER(C.c_float)
    sig = [C.c_int] * 5 + [C.c_float] * 4 + [C.c_int] * 5 + [F] * 14 + [C.c_char_p]
    for sym in ("modeling3d", "_Z10modeling3diiiiiffffiiiiiPfS_S_S_S_S_S_S_S_S_S_S_S_S_PKc"):
        fn = getattr(lib, sym)
        fn.restype = None
        fn.argtypes = sig
        rec = np.zeros((s.nt, s.hx, s.hy), np.float32)
        arrs = [np.ascontiguousarray(a, np.float32).reshape(-1) for a in
                (s.vp, s.eps, s.dele, s.source, s.wavelet, rec, s.dlr, s.ddlr, s.dtb, s.ddtb, s.dfb, s.ddfb, s.c,
                 s.c2)]
        sx, sy, _ = s.src_cell
        print(f"void entry point {sym[:16]}", flush=True)
        fn(s.nx, s.ny, s.nz, s.nt, 1, s.dx, s.dy, s.dz, s.dt, s.pml, 0, sx, sy, s.sz, *[fptr(a) for a in arrs], None)
        got = arrs[5].reshape(s.nt, s.hx, s.hy)
        assert rel_l2(got, orec) <= TOL, sym


def test_dropin_returns_the_record_of_a_blown_up_run():
    """A CFL violation: the status-returning entry point reports QP3D_ERR_NONFINITE and still
    writes the record, which holds the non-finite values as the reference's would."""
    s = qp3d.config_setup(hx=21, hy=21, nz0=21, pml=8, nt=400, seed=0, dt=0.004)
    rec = np.zeros((s.nt, s.hx, s.hy), np.float32)
    with pytest.raises(qp3d.Qp3dError) as ei:
        qp3d.modeling3d(s, record=rec)
    assert ei.value.code == -5
    assert not np.all(np.isfinite(rec)) and np.any(rec[:50] != 0)
