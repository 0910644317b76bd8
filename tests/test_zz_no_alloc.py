"""The allocation-free entry point qp3d_modeling3d_noalloc (device arrays in, device record out,
all other device memory from the library's fixed scratch array) against qp3d_modeling3d, the
entry point it is a variant of: two calls in a row, the second on a smaller grid (the scratch
array keeps the first call's contents, which the second must not see), bit for bit."""
import ctypes as C
import sys

import numpy as np
import pytest

import qp3d
from qp3d._lib import QP3D_ERR_ARG, fptr

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _hip():
    qp3d.load()
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    for f in (hip.hipMalloc, hip.hipMemcpy, hip.hipFree, hip.hipStreamCreate, hip.hipStreamDestroy):
        f.restype = C.c_int
    return hip


def _args(s):
    return [s.vp, s.eps, s.dele, s.source, s.wavelet, None, s.dlr, s.ddlr, s.dtb, s.ddtb, s.dfb, s.ddfb, s.c, s.c2]


def noalloc(hip, s, stream):
    """Runs one shot through qp3d_modeling3d_noalloc with device copies of the inputs."""
    lib = qp3d.load()
    bufs = []
    try:
        ptrs = []
        rec_bytes = 4 * s.nt * s.hx * s.hy
        for a in _args(s):
            p = C.c_void_p()
            if a is None:
                assert hip.hipMalloc(C.byref(p), rec_bytes) == 0
            else:
                a = np.ascontiguousarray(a, np.float32).reshape(-1)
                assert hip.hipMalloc(C.byref(p), a.nbytes) == 0
                assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
            bufs.append(p)
            ptrs.append(C.cast(p, C.POINTER(C.c_float)))
        sx, sy, _ = s.src_cell
        rc = lib.qp3d_modeling3d_noalloc(s.nx, s.ny, s.nz, s.nt, 1, s.dx, s.dy, s.dz, s.dt, s.pml, 0, sx, sy, s.sz,
                                         *ptrs, None, stream)
        assert rc == 0, lib.qp3d_last_error()
        out = np.empty((s.nt, s.hx, s.hy), np.float32)
        assert hip.hipMemcpy(out.ctypes.data, bufs[5], rec_bytes, 2) == 0
        return out
    finally:
        for p in bufs:
            hip.hipFree(p)


def test_noalloc_matches_modeling3d_on_two_calls_in_a_row():
    hip = _hip()
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    try:
        cases = [qp3d.config_setup(hx=45, hy=41, nz0=39, pml=12, nt=120, seed=3, src_depth=3),
                 qp3d.config_setup(hx=23, hy=27, nz0=21, pml=9, nt=90, seed=5, src_depth=2)]
        for i, s in enumerate(cases):
            say(f"noalloc call {i}: {s.nx}x{s.ny}x{s.nz}, {s.nt} steps")
            got = noalloc(hip, s, stream)
            ref = qp3d.modeling3d(s)
            assert np.any(ref != 0)
            assert np.array_equal(got, ref), i
    finally:
        hip.hipStreamDestroy(stream)


def test_noalloc_rejects_a_grid_beyond_its_scratch_and_host_arrays():
    """A grid larger than the scratch array and host-memory arrays are refused before any
    device work (QP3D_ERR_ARG with a message)."""
    hip = _hip()
    lib = qp3d.load()
    s = qp3d.config_setup(hx=23, hy=21, nz0=19, pml=8, nt=10, seed=1)
    say("noalloc: host arrays")
    host = [np.ascontiguousarray(a, np.float32).reshape(-1) if a is not None else
            np.zeros(s.nt * s.hx * s.hy, np.float32) for a in _args(s)]
    sx, sy, _ = s.src_cell
    rc = lib.qp3d_modeling3d_noalloc(s.nx, s.ny, s.nz, s.nt, 1, s.dx, s.dy, s.dz, s.dt, s.pml, 0, sx, sy, s.sz,
                                     *[fptr(a) for a in host], None, None)
    assert rc == QP3D_ERR_ARG and b"device memory" in lib.qp3d_last_error()
    say("noalloc: grid beyond the scratch array")
    big = (200, 200, 200)  # 8e6 cells, ~350 MB of state: beyond the scratch array
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), 4 * 200 ** 3) == 0
    try:
        f = C.cast(p, C.POINTER(C.c_float))
        rc = lib.qp3d_modeling3d_noalloc(*big, 10, 1, 12.5, 12.5, 12.5, 0.001, 20, 0, 0, 0, 30, *([f] * 14), None,
                                         None)
        assert rc == QP3D_ERR_ARG and b"scratch" in lib.qp3d_last_error()
    finally:
        hip.hipFree(p)
