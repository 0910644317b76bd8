"""Host setup math of the library against the reference's own setup functions (golden vectors
written from CPU_function.c, and oracle/_ref directly when built), the synthetic model rule,
model-file reads and the C-ABI surface (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import qp3d
from qp3d._lib import fptr, load, QP3D_ERR_ARG, QP3D_ERR_IO
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "ref_setup.npz"))


def test_fd_tables_match_reference(gold):
    c, c2 = qp3d.fd_tables()
    assert np.array_equal(bits(c), bits(gold["c"]))
    assert np.array_equal(bits(c2), bits(gold["c2"]))


def test_pml_profiles_match_reference(gold):
    lib = load()
    for n in range(4):
        pml, vpmax, dx, dy, dz = gold[f"pml{n}_args"]
        pml = int(pml)
        tabs = [np.zeros(pml, np.float32) for _ in range(6)]
        lib.qp3d_pmlcoff(pml, float(vpmax), float(dx), float(dy), float(dz), *[fptr(t) for t in tabs])
        for name, t in zip(["dlr", "ddlr", "dtb", "ddtb", "dfb", "ddfb"], tabs):
            assert np.array_equal(bits(t), bits(gold[f"pml{n}_{name}"])), (n, name)


def test_wavelets_match_reference(gold):
    lib = load()
    for n in range(3):
        fm, dt, nt, t0 = gold[f"wav{n}_args"]
        wav = np.zeros(int(nt), np.float32)
        lib.qp3d_source(0, 0, 0, float(fm), 1.0, 1.0, float(dt), 10.0, 10.0, 10.0, float(t0), int(nt), 1, 1, 1,
                        None, fptr(wav))
        assert np.array_equal(bits(wav), bits(gold[f"wav{n}"])), n


def test_spatial_source_matches_reference(gold):
    lib = load()
    a = gold["src_args"]
    nx, ny, nz = int(a[12]), int(a[13]), int(a[14])
    src = np.zeros(nx * ny * nz, np.float32)
    lib.qp3d_source(int(a[0]), int(a[1]), int(a[2]), *[float(v) for v in a[3:11]], int(a[11]), nx, ny, nz,
                    fptr(src), None)
    assert np.array_equal(bits(src), bits(gold["src"]))
    assert np.count_nonzero(src) > 1


def test_extend_model_matches_reference(gold):
    lib = load()
    hx, hy, hz, pml = [int(v) for v in gold["ext_dims"]]
    out = np.zeros_like(gold["ext_out"])
    lib.qp3d_extend_model(fptr(np.ascontiguousarray(gold["ext_in"])), fptr(out), hx + 2 * pml, hy + 2 * pml,
                          hz + 2 * pml, pml)
    assert np.array_equal(bits(out), bits(gold["ext_out"]))


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref needs the reference checkout")
def test_setup_matches_reference_live():
    R = O.ref_lib()
    lib = load()
    rng = np.random.default_rng(3)
    for pml, vpmax, h in ((6, 3000.0, 7.5), (17, 4123.4, 12.5), (40, 5100.0, 20.0)):
        a = [np.zeros(pml, np.float32) for _ in range(12)]
        R.ref_pmlcoff(pml, vpmax, h, h * 1.1, h * 0.9, *[O._fp(t) for t in a[:6]])
        lib.qp3d_pmlcoff(pml, vpmax, h, h * 1.1, h * 0.9, *[fptr(t) for t in a[6:]])
        for x, y in zip(a[:6], a[6:]):
            assert np.array_equal(bits(x), bits(y))
    nx, ny, nz = 17, 15, 13
    s1, s2 = np.zeros(nx * ny * nz, np.float32), np.zeros(nx * ny * nz, np.float32)
    w1, w2 = np.zeros(700, np.float32), np.zeros(700, np.float32)
    args = (8, 7, 6, 17.5, 2.0, 0.05, 0.0005, 9.0, 11.0, 7.0, float(np.float32(1 / 17.5)), 700, nx, ny, nz)
    R.ref_source(*args, O._fp(s1), O._fp(w1))
    lib.qp3d_source(*args, fptr(s2), fptr(w2))
    assert np.array_equal(bits(s1), bits(s2)) and np.array_equal(bits(w1), bits(w2))
    inner = rng.standard_normal(5 * 6 * 7).astype(np.float32)
    e1 = np.zeros(15 * 16 * 17, np.float32)
    e2 = np.zeros_like(e1)
    R.ref_extend_model(O._fp(inner), O._fp(e1), 15, 16, 17, 5)
    lib.qp3d_extend_model(fptr(inner), fptr(e2), 15, 16, 17, 5)
    assert np.array_equal(e1, e2)


# ---- synthetic model rule -------------------------------------------------------------

def test_genmodel_layers_bounds_and_determinism():
    vp, ep, de = qp3d.genmodel(20, 18, 40, seed=0)
    assert vp.shape == (18, 20, 40)  # [y][x][z] file layout
    base_v = np.array([2000, 2500, 3000, 3500], np.float64)
    base_e = np.array([0, 0.10, 0.15, 0.20])
    layer = (np.arange(40) * 4) // 40
    rel = vp / base_v[layer][None, None, :] - 1
    assert np.abs(rel).max() <= 0.03 + 1e-6 and np.abs(rel).max() > 0.02
    assert np.all(ep[:, :, layer == 0] == 0) and np.all(de[:, :, layer == 0] == 0)
    rel_e = ep[:, :, layer > 0] / base_e[layer[layer > 0]][None, None, :] - 1
    assert np.abs(rel_e).max() <= 0.03 + 1e-6
    assert np.all(de <= ep)
    vp2, ep2, de2 = qp3d.genmodel(20, 18, 40, seed=0)
    assert np.array_equal(vp, vp2) and np.array_equal(ep, ep2) and np.array_equal(de, de2)
    vp3, _, _ = qp3d.genmodel(20, 18, 40, seed=1)
    assert not np.array_equal(vp, vp3)
    # the three perturbations are independent draws
    gv = vp / base_v[layer][None, None, :] - 1
    ge = np.zeros_like(gv)
    ge[:, :, layer > 0] = rel_e
    assert abs(np.corrcoef(gv[:, :, layer > 0].ravel(), ge[:, :, layer > 0].ravel())[0, 1]) < 0.5


def test_genmodel_perturbation_is_smooth():
    vp, _, _ = qp3d.genmodel(48, 48, 48, seed=2)
    g = vp[:, :, :12] / 2000.0 - 1  # top layer, constant base
    d1 = np.abs(np.diff(g, axis=1)).mean()
    assert d1 < 0.3 * np.abs(g).mean()  # ~0.14 for sigma = 5, ~1.1 for white noise


def test_readmodel3d_window_and_errors(tmp_path):
    lib = load()
    vny, vnx, vnz = 9, 8, 7
    data = np.arange(vny * vnx * vnz, dtype=np.float32)
    f = tmp_path / "m.dat"
    data.tofile(f)
    ny, nx, nz, fny, fnx = 4, 3, 7, 2, 5
    a = np.zeros(ny * nx * nz, np.float32)
    assert lib.qp3d_readmodel3d(str(f).encode(), fptr(a), ny, nx, nz, fny, fnx, 0, vny, vnx, vnz) == 0
    expect = data.reshape(vny, vnx, vnz)[fny:fny + ny, fnx:fnx + nx, :].transpose(2, 1, 0)
    assert np.array_equal(a.reshape(nz, nx, ny), expect)
    if O.ref_available():
        b = np.zeros_like(a)
        O.ref_lib().ref_readmodel3d(str(f).encode(), O._fp(b), ny, nx, nz, fny, fnx, 0, vny, vnx, vnz)
        assert np.array_equal(a, b)
    # window outside the model (Main.c:187 with sx - (hx-1)/2 < 0) and a missing file
    assert lib.qp3d_readmodel3d(str(f).encode(), fptr(a), ny, nx, nz, -1, 0, 0, vny, vnx, vnz) == QP3D_ERR_ARG
    assert lib.qp3d_readmodel3d(str(f).encode(), fptr(a), ny, 4, nz, 0, 5, 0, vny, vnx, vnz) == QP3D_ERR_ARG
    assert lib.qp3d_readmodel3d(b"/nonexistent/vp.dat", fptr(a), ny, nx, nz, 0, 0, 0, vny, vnx, vnz) == QP3D_ERR_IO
    short = tmp_path / "short.dat"
    data[:100].tofile(short)
    assert lib.qp3d_readmodel3d(str(short).encode(), fptr(a), ny, nx, nz, 0, 0, 0, vny, vnx, vnz) == QP3D_ERR_IO
    assert "shorter" in lib.qp3d_last_error().decode()


# ---- C ABI surface ----------------------------------------------------------------------

def header_symbols():
    text = open(os.path.join(REPO, "include", "qp3d.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(modeling3d|qp3d_[a-z0-9_]+)\s*\(", text)))


def test_tile_count_matches_the_header():
    text = open(os.path.join(REPO, "include", "qp3d.h")).read()
    assert int(re.search(r"#define QP3D_NTILES (\d+)", text).group(1)) == qp3d.NTILES


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(qp3d.LIB_PATH)
    syms = header_symbols()
    assert "modeling3d" in syms and "qp3d_plan_run" in syms
    for name in syms:
        assert hasattr(lib, name), name
    assert set(syms) <= set(qp3d.EXPORTED)


def test_plan_create_rejects_bad_config_without_touching_the_device():
    lib = load()
    from qp3d._lib import Config
    h = C.c_void_p()
    bad = [Config(40, 40, 40, 5, 10, 10, 10, 10, 0.001, 5, 0, 0, 0, 0, 0, 0),     # pml < 6
           Config(20, 40, 40, 10, 10, 10, 10, 10, 0.001, 10, 0, 0, 0, 0, 0, 0),   # nx <= 2 pml
           Config(40, 40, 40, 10, 10, 10, 10, 10, 0.0, 10, 0, 0, 0, 0, 0, 0),     # dt = 0
           Config(40, 40, 40, 10, 10, 10, 10, 10, 0.001, 40, 0, 0, 0, 0, 0, 0),   # sz outside
           Config(40, 40, 40, 10, 10, 10, 10, 10, 0.001, 10, 0, 0, 1, 0, 0, 0),   # snapflag, ntsnap 0
           Config(40, 40, 40, 10, 10, 10, 10, 10, 0.001, 10, 7, 0, 0, 0, 0, 0),   # unknown storage
           Config(40, 40, 40, 10, 10, 10, 10, 10, 0.001, 10, 0, 0, 0, 0, 0, 0, snap_s=1),  # S field, no snapshots
           Config(40, 40, 40, 10, 10, 10, 10, 10, 0.001, 10, 0, 0, 0, 0, 0, 0, gmax_mode=3)]  # unknown G mode
    for cfg in bad:
        assert lib.qp3d_plan_create(C.byref(cfg), C.byref(h)) == QP3D_ERR_ARG
        assert lib.qp3d_last_error().decode()


# The reference declares modeling3d without extern "C" and compiles its driver with g++
# (3dPureqPModeling_single_shot.h:4-6, Makefile:3,11-13): a driver built that way refers to the
# C++-mangled symbol, which the library exports beside the C one.
CXX_CALLER = r"""
void modeling3d(int nx, int ny, int nz, int nt, int ntsnap, float dx, float dy, float dz, float dt, int pml,
                int snapflag, int sx, int sy, int sz, float *vp, float *epsilon, float *delta, float *source,
                float *wavelet, float *record, float *dlr, float *ddlr, float *dtb, float *ddtb, float *dfb,
                float *ddfb, float *c, float *c2, const char *snap_file);
int main(int argc, char **argv)
{
  if (argc > 99) modeling3d(0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0);
  return 0;
}
"""


def _cxx_header_declaration():
    text = open(os.path.join(REPO, "include", "qp3d_cxx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.search(r"void modeling3d\(.*?\);", text, flags=re.S).group(0)


def test_cxx_linkage_modeling3d_links_from_a_gpp_driver(tmp_path):
    """A g++-compiled caller of the reference's declaration links against libqp3d.so with no
    header edit; the declaration in include/qp3d_cxx.h is that same one, and the C++ entry point
    is exported next to the C one."""
    import shutil
    import subprocess
    lib_dir = os.path.dirname(qp3d.LIB_PATH)
    src = tmp_path / "caller.cpp"
    src.write_text(CXX_CALLER)
    gxx = shutil.which("g++")
    assert gxx
    r = subprocess.run([gxx, "-o", str(tmp_path / "caller"), str(src), "-L" + lib_dir, "-lqp3d",
                        "-Wl,-rpath," + lib_dir], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    decl = " ".join(_cxx_header_declaration().split())
    want = " ".join(re.search(r"void modeling3d\(.*?\);", CXX_CALLER, flags=re.S).group(0).split())
    assert decl == want
    nm = subprocess.run(["nm", "-D", "--defined-only", qp3d.LIB_PATH], capture_output=True, text=True).stdout
    assert re.search(r" T modeling3d$", nm, flags=re.M)
    assert re.search(r" T _Z10modeling3diiiiiffffiiiiiPfS_S_S_S_S_S_S_S_S_S_S_S_S_PKc$", nm, flags=re.M)


@pytest.mark.skipif(not os.path.exists(os.path.join(REPO, "oracle", "_ref", "Main.o")),
                    reason="the reference's driver object is built only where the reference checkout is")
def test_reference_driver_object_links_against_the_library(tmp_path):
    """The reference's Main.c, compiled unchanged with g++ (oracle/Makefile), links against
    libqp3d.so in place of modeling3D.o: its only unresolved solver symbol is the C++ one."""
    import subprocess
    ref = os.path.join(REPO, "oracle", "_ref")
    objs = [os.path.join(ref, f) for f in ("Main.o", "array_new.o", "read_write.o", "CPU_function.o")]
    und = subprocess.run(["nm", "-u", objs[0]], capture_output=True, text=True).stdout
    assert "_Z10modeling3diiiiiffffiiiiiPfS_S_S_S_S_S_S_S_S_S_S_S_S_PKc" in und
    lib_dir = os.path.dirname(qp3d.LIB_PATH)
    r = subprocess.run(["g++", "-o", str(tmp_path / "ref_main"), *objs, "-L" + lib_dir, "-lqp3d",
                        "-Wl,-rpath," + lib_dir, "-lm"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
