"""ctypes wrapper of the CPU oracle (liboracle.so) and of oracle/_ref/libqp3d_ref.so.

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline as the checker / the timed CPU reference, never by the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libqp3d_ref.so")

FIELDS = {"u1": 0, "u2": 1, "u3": 2, "ux": 3, "uy": 4, "uz": 5, "S": 6}
FACES = ["L", "R", "F", "Bk", "T", "Bo"]
AUX = ["w11", "w12", "w13", "w21", "w31", "w32", "w33", "p1", "p2", "p3"]

_F = C.POINTER(C.c_float)
_oracle = None
_ref = None


def default_threads():
    """Threads for the oracle: OMP_NUM_THREADS when set (the host's CPU share), else all cores."""
    v = os.environ.get("OMP_NUM_THREADS")
    return int(v) if v and v.isdigit() and int(v) > 0 else (os.cpu_count() or 1)


def build(ref=False):
    subprocess.run(["make", "-s", "-C", HERE] + (["ref"] if ref else []), check=True)


def _load_oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            build()
        lib = C.CDLL(ORACLE_SO)
        for sfx, R in (("f32", C.c_float), ("f64", C.c_double)):
            RP = C.POINTER(R)
            f = getattr(lib, f"oracle_create_{sfx}")
            f.restype = C.c_void_p
            f.argtypes = [C.c_int] * 4 + [C.c_float] * 4 + [_F] * 12 + [C.c_int]
            f = getattr(lib, f"oracle_step_{sfx}")
            f.restype = None
            f.argtypes = [C.c_void_p, C.c_float]
            f = getattr(lib, f"oracle_step_planted_G_{sfx}")
            f.restype = None
            f.argtypes = [C.c_void_p, C.c_float, C.c_double]
            f = getattr(lib, f"oracle_field_{sfx}")
            f.restype = RP
            f.argtypes = [C.c_void_p, C.c_int]
            f = getattr(lib, f"oracle_aux_{sfx}")
            f.restype = RP
            f.argtypes = [C.c_void_p, C.c_int, C.c_int]
            f = getattr(lib, f"oracle_aux_len_{sfx}")
            f.restype = C.c_long
            f.argtypes = [C.c_void_p, C.c_int]
            f = getattr(lib, f"oracle_last_G_{sfx}")
            f.restype = C.c_double
            f.argtypes = [C.c_void_p]
            f = getattr(lib, f"oracle_set_decisions_{sfx}")
            f.restype = None
            f.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
            f = getattr(lib, f"oracle_run_record_{sfx}")
            f.restype = None
            f.argtypes = [C.c_void_p, _F, C.c_int, C.c_int, RP]
            f = getattr(lib, f"oracle_destroy_{sfx}")
            f.restype = None
            f.argtypes = [C.c_void_p]
            f = getattr(lib, f"oracle_modeling3d_{sfx}")
            f.restype = C.c_int
            f.argtypes = [C.c_int] * 5 + [C.c_float] * 4 + [C.c_int] * 3 + [_F] * 5 + [RP] + [_F] * 8 + \
                [C.c_char_p, C.c_int]
        _oracle = lib
    return _oracle


def ref_available():
    return os.path.exists(REF_SO)


def _load_ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(f"{REF_SO} not built (make -C oracle ref, needs the reference checkout)")
        lib = C.CDLL(REF_SO)
        lib.ref_create.restype = C.c_void_p
        lib.ref_create.argtypes = [C.c_int] * 4 + [C.c_float] * 4 + [_F] * 12
        lib.ref_step.restype = None
        lib.ref_step.argtypes = [C.c_void_p, C.c_float]
        lib.ref_field.restype = _F
        lib.ref_field.argtypes = [C.c_void_p, C.c_int]
        lib.ref_aux.restype = _F
        lib.ref_aux.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.ref_aux_len.restype = C.c_long
        lib.ref_aux_len.argtypes = [C.c_void_p, C.c_int]
        lib.ref_last_G.restype = C.c_double
        lib.ref_last_G.argtypes = [C.c_void_p]
        lib.ref_destroy.restype = None
        lib.ref_destroy.argtypes = [C.c_void_p]
        for n in ("ref_fdcoff1", "ref_fdcoff2"):
            getattr(lib, n).argtypes = [_F]
        lib.ref_pmlcoff.argtypes = [C.c_int] + [C.c_float] * 4 + [_F] * 6
        lib.ref_source.argtypes = [C.c_int] * 3 + [C.c_float] * 8 + [C.c_int] * 4 + [_F, _F]
        lib.ref_extend_model.argtypes = [_F, _F] + [C.c_int] * 4
        lib.ref_absmax.restype = C.c_float
        lib.ref_absmax.argtypes = [_F, C.c_int, C.c_int, C.c_int]
        lib.ref_readmodel3d.argtypes = [C.c_char_p, _F] + [C.c_int] * 9
        _ref = lib
    return _ref


def ref_lib():
    return _load_ref()


def _fp(a):
    return a.ctypes.data_as(_F)


def _inputs(s):
    keep = [np.ascontiguousarray(a, np.float32).reshape(-1) for a in
            (s.vp, s.eps, s.dele, s.source, s.dlr, s.ddlr, s.dtb, s.ddtb, s.dfb, s.ddfb, s.c, s.c2)]
    return keep


class _StateBase:
    """Common field / aux access over a state handle."""
    dtype = np.float32

    def _view(self, ptr, n):
        return np.ctypeslib.as_array(ptr, shape=(n,))

    def field(self, name):
        s = self.s
        n = s.nx * s.ny * s.nz
        return self._view(self._field(FIELDS[name]), n).reshape(s.nz, s.nx, s.ny)

    def aux(self, face, arr):
        f = FACES.index(face)
        return self._view(self._aux(f, AUX.index(arr)), self._aux_len(f))

    def set_field(self, name, a):
        self.field(name)[...] = a

    def set_aux(self, face, arr, a):
        self.aux(face, arr)[...] = a

    def import_state(self, u_new, u_old, w1_new, w1_old, w3_new, w3_old, p_new, p_old, w21):
        """Load a mid-run state given in the full-grid layout (PML state held at the cells each
        face owns: x faces first, then y, then z) into the per-face arrays, as it stands after
        the reference's exchange: u2 = u3 = newest, u1 = older, w12/w32/p2 = newest,
        w11/w31/p1 = older, w21 = accumulator."""
        s = self.s
        nx, ny, nz, pml = s.nx, s.ny, s.nz, s.pml
        self.set_field("u2", u_new)
        self.set_field("u3", u_new)
        self.set_field("u1", u_old)
        xs = slice(pml, nx - pml)
        ys = slice(pml, ny - pml)

        def face_view(a, face):
            if face == "L":
                return a[:, :pml, :].copy()
            if face == "R":
                return a[:, nx - pml:, :].copy()
            if face in ("F", "Bk"):
                v = a[:, :, :pml] if face == "F" else a[:, :, ny - pml:]
                v = v.copy()
                v[:, :pml, :] = 0
                v[:, nx - pml:, :] = 0
                return v
            v = (a[:pml] if face == "T" else a[nz - pml:]).copy()
            m = np.ones(v.shape[1:], bool)
            m[xs, ys] = False
            v[:, m] = 0
            return v

        pairs = [("w12", w1_new), ("w11", w1_old), ("w32", w3_new), ("w31", w3_old), ("p2", p_new),
                 ("p1", p_old), ("w21", w21)]
        for face in FACES:
            for arr, src in pairs:
                self.set_aux(face, arr, face_view(np.asarray(src), face).reshape(-1))


class OracleState(_StateBase):
    """The oracle's state machine: step(w) = grad, max, S, update, source, exchange."""

    def __init__(self, s, precision="f32", nthreads=None):
        lib = _load_oracle()
        self.s = s
        self.sfx = precision
        self.dtype = np.float32 if precision == "f32" else np.float64
        self._keep = _inputs(s)
        nthreads = nthreads or default_threads()
        self.lib = lib
        self.h = getattr(lib, f"oracle_create_{precision}")(
            s.nx, s.ny, s.nz, s.pml, s.dx, s.dy, s.dz, s.dt, *[_fp(a) for a in self._keep], int(nthreads))
        if not self.h:
            raise MemoryError("oracle_create failed")

    def _field(self, i):
        return getattr(self.lib, f"oracle_field_{self.sfx}")(self.h, i)

    def _aux(self, f, a):
        return getattr(self.lib, f"oracle_aux_{self.sfx}")(self.h, f, a)

    def _aux_len(self, f):
        return getattr(self.lib, f"oracle_aux_len_{self.sfx}")(self.h, f)

    def step(self, w):
        getattr(self.lib, f"oracle_step_{self.sfx}")(self.h, float(w))

    def step_planted_G(self, w, G_threshold):
        """A planted mistake for the tests: one step whose scalar operator takes its
        threshold 0.01*G from G_threshold instead of this step's G."""
        getattr(self.lib, f"oracle_step_planted_G_{self.sfx}")(self.h, float(w), float(G_threshold))

    @property
    def G(self):
        return getattr(self.lib, f"oracle_last_G_{self.sfx}")(self.h)

    def set_decisions(self, bits, mode):
        """Test hook: mode 1 records each following step's per-cell decision |grad u| > 0.01*G
        into `bits` (uint8, nt x ceil(N/8)), mode 2 replays them in place of the comparison (the
        linearisation point of the adjoint: threshold and decisions held), 0 turns it off."""
        self._bits = bits
        ptr = None if bits is None else bits.ctypes.data_as(C.c_void_p)
        getattr(self.lib, f"oracle_set_decisions_{self.sfx}")(self.h, ptr, int(mode))

    def run_record(self, nt=None, wavelet=None):
        """nt steps (the setup's wavelet by default); returns the record [nt][hx][hy]."""
        s = self.s
        nt = s.nt if nt is None else nt
        w = np.ascontiguousarray(s.wavelet if wavelet is None else wavelet, np.float32)
        rec = np.zeros((nt, s.hx, s.hy), self.dtype)
        R = C.c_float if self.sfx == "f32" else C.c_double
        getattr(self.lib, f"oracle_run_record_{self.sfx}")(self.h, _fp(w), int(nt), int(s.sz),
                                                           rec.ctypes.data_as(C.POINTER(R)))
        return rec

    def close(self):
        if self.h:
            getattr(self.lib, f"oracle_destroy_{self.sfx}")(self.h)
            self.h = None

    def __del__(self):
        self.close()


class RefState(_StateBase):
    """The reference's own CPU kernels (oracle/_ref) driven in the same step order."""

    def __init__(self, s):
        lib = _load_ref()
        self.s = s
        self.lib = lib
        self._keep = _inputs(s)
        self.h = lib.ref_create(s.nx, s.ny, s.nz, s.pml, s.dx, s.dy, s.dz, s.dt, *[_fp(a) for a in self._keep])

    def _field(self, i):
        return self.lib.ref_field(self.h, i)

    def _aux(self, f, a):
        return self.lib.ref_aux(self.h, f, a)

    def _aux_len(self, f):
        return self.lib.ref_aux_len(self.h, f)

    def step(self, w):
        self.lib.ref_step(self.h, float(w))

    @property
    def G(self):
        return self.lib.ref_last_G(self.h)

    def close(self):
        if self.h:
            self.lib.ref_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def decision_bits(s, nt=None):
    """A zeroed buffer for set_decisions: nt steps of one bit per cell."""
    n = s.nx * s.ny * s.nz
    return np.zeros(((s.nt if nt is None else nt), (n + 7) // 8), np.uint8)


def record_frozen(s, bits, precision="f64", nthreads=None):
    """Record of a run whose per-cell threshold decisions are replayed from `bits`."""
    o = OracleState(s, precision, nthreads=nthreads)
    try:
        o.set_decisions(bits, 2)
        return o.run_record()
    finally:
        o.close()


def modeling3d(s, precision="f32", nthreads=None, snapflag=0, ntsnap=1, snap_file=None, nt=None):
    """The oracle's modeling3d (modeling3D.c:17-153 order). Returns record [nt][hx][hy]."""
    lib = _load_oracle()
    nt = s.nt if nt is None else nt
    dt_ = np.float32 if precision == "f32" else np.float64
    rec = np.zeros((nt, s.hx, s.hy), dt_)
    k = _inputs(s)
    wav = np.ascontiguousarray(s.wavelet, np.float32)
    R = C.c_float if precision == "f32" else C.c_double
    rc = getattr(lib, f"oracle_modeling3d_{precision}")(
        s.nx, s.ny, s.nz, nt, int(ntsnap), s.dx, s.dy, s.dz, s.dt, s.pml, int(snapflag), s.sz,
        _fp(k[0]), _fp(k[1]), _fp(k[2]), _fp(k[3]), _fp(wav), rec.ctypes.data_as(C.POINTER(R)),
        *[_fp(a) for a in k[4:]], snap_file.encode() if snap_file else None, int(nthreads or default_threads()))
    if rc:
        raise MemoryError("oracle_modeling3d failed")
    return rec
