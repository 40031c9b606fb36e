"""ctypes binding of libartra_hip.so (declarations in include/artra_hip.h).

Names follow the solver's own vocabulary: time levels TO/TN/TM = 0/1/2,
node flags did/fid/lid/gst, interior range ``pin`` (ns[PIN] of the
reference), conservative components (rho, rho u, rho v, rho w, rho eT).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libartra_hip.so"
TO, TN, TM = 0, 1, 2

_lib = None


MLP_PARAMS = 1633


class MlpTraining(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_int), ("batch", ctypes.c_int), ("lr", ctypes.c_float), ("tau", ctypes.c_float),
                ("wpos", ctypes.c_float)]


class ArtraError(RuntimeError):
    """A libartra_hip call returned a non-zero code."""

    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


def library_path():
    return os.path.join(HERE, _LIB_NAME)


def library():
    """Load libartra_hip.so (built in-tree by artracfd_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: run __graft_entry__.build() or python -m artracfd_amd.build")
    lib = ctypes.CDLL(path)
    vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    lp = ctypes.POINTER(ctypes.c_long)
    fp = ctypes.POINTER(ctypes.c_float)
    sig = {
        "artra_hip_attach": (ctypes.c_int, [vp, vp, ctypes.POINTER(vp)]),
        "artra_hip_detach": (ctypes.c_int, [vp]),
        "artra_hip_treat_boundary": (ctypes.c_int, [vp, ctypes.c_int]),
        "artra_hip_compute_dt": (ctypes.c_int, [vp, vp, dp]),
        "artra_hip_evolve": (ctypes.c_int, [vp, ctypes.c_double]),
        "artra_hip_march": (ctypes.c_int, [vp, vp, ctypes.c_int, ip, dp]),
        "artra_hip_march_noalloc": (ctypes.c_int, [vp, vp, ctypes.c_int, ip, dp]),
        "artra_hip_surface_force": (ctypes.c_int, [vp]),
        "artra_hip_write_probe_data": (ctypes.c_int, [vp, vp, ctypes.c_int]),
        "artra_hip_write_space_data": (ctypes.c_int, [vp, vp]),
        "artra_hip_solve": (ctypes.c_int, [vp, vp]),
        "artra_hip_write_restart": (ctypes.c_int, [vp, vp, ctypes.c_char_p]),
        "artra_hip_read_restart": (ctypes.c_int, [vp, vp, ctypes.c_char_p]),
        "artra_hip_download": (ctypes.c_int, [vp, ctypes.c_int]),
        "artra_hip_upload": (ctypes.c_int, [vp, ctypes.c_int]),
        "artra_hip_read_field": (ctypes.c_int, [vp, ctypes.c_int, dp]),
        "artra_hip_write_field": (ctypes.c_int, [vp, ctypes.c_int, dp]),
        "artra_hip_synchronize": (ctypes.c_int, [vp]),
        "artra_hip_set_timing": (ctypes.c_int, [vp, ctypes.c_int]),
        "artra_hip_get_timing": (ctypes.c_int, [vp, dp, lp]),
        "artra_hip_set_tuning": (ctypes.c_int, [vp, ctypes.c_int]),
        "artra_hip_set_arithmetic": (ctypes.c_int, [vp, ctypes.c_int]),
        "artra_hip_set_fill_options": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "artra_hip_set_graph": (ctypes.c_int, [vp, ctypes.c_int]),
        "artra_hip_set_fused": (ctypes.c_int, [vp, ctypes.c_int]),
        "artra_hip_set_indicator_options": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_float]),
        "artra_hip_fused_axes": (ctypes.c_int, [vp, ip]),
        "artra_hip_error": (ctypes.c_char_p, []),
        "artra_hip_copy_probe": (ctypes.c_int, [ctypes.c_long, ctypes.c_int, dp]),
        "artra_case_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(vp)]),
        "artra_case_close": (None, [vp]),
        "artra_case_space": (vp, [vp]),
        "artra_case_model": (vp, [vp]),
        "artra_case_time": (vp, [vp]),
        "artra_case_grid": (ctypes.c_int, [vp, ip, ip, ip]),
        "artra_case_get_field": (ctypes.c_int, [vp, ctypes.c_int, dp]),
        "artra_case_set_field": (ctypes.c_int, [vp, ctypes.c_int, dp]),
        "artra_case_get_flags": (ctypes.c_int, [vp, ip]),
        "artra_case_clock": (ctypes.c_int, [vp, dp, ip]),
        "artra_case_forces": (ctypes.c_int, [vp, dp]),
        "artra_case_bodies": (ctypes.c_int, [vp, dp]),
        "artra_calc_expression": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_double, ctypes.c_double,
                                                 ctypes.c_double, dp]),
        "artra_hip_face_count": (ctypes.c_int, [vp, ctypes.c_int, lp]),
        "artra_hip_sample_faces": (ctypes.c_int, [vp, ctypes.c_int, fp, fp]),
        "artra_hip_set_indicator": (ctypes.c_int, [vp, fp, ctypes.c_float]),
        "artra_hip_hybrid_counts": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_int]),
        "artra_mlp_train": (ctypes.c_int, [fp, fp, fp, ctypes.c_long, ctypes.POINTER(MlpTraining), fp]),
        "artra_mlp_gradient": (ctypes.c_int, [fp, fp, fp, ctypes.c_long, ctypes.c_float, ctypes.c_float, fp, fp]),
        "artra_mlp_logits": (ctypes.c_int, [fp, fp, ctypes.c_long, fp]),
        "EvolveFluidDynamics": (None, [ctypes.c_double, vp, vp]),
        "TreatBoundary": (None, [ctypes.c_int, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(code):
    if code != 0:
        raise ArtraError(code, library().artra_hip_error().decode(errors="replace"))


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def mlp_train(params, features, labels, epochs=3, batch=4096, lr=3e-3, tau=0.1, wpos=4.0):
    """Train the indicator on the device; returns (params, loss per epoch)."""
    p = np.ascontiguousarray(params, dtype=np.float32).copy()
    f = np.ascontiguousarray(features, dtype=np.float32)
    y = np.ascontiguousarray(labels, dtype=np.float32)
    hist = np.zeros(epochs, dtype=np.float32)
    opt = MlpTraining(epochs, batch, lr, tau, wpos)
    _check(library().artra_mlp_train(_fptr(p), _fptr(f), _fptr(y), len(y), ctypes.byref(opt), _fptr(hist)))
    return p, hist


def mlp_gradient(params, features, labels, tau, wpos):
    p = np.ascontiguousarray(params, dtype=np.float32)
    f = np.ascontiguousarray(features, dtype=np.float32)
    y = np.ascontiguousarray(labels, dtype=np.float32)
    g = np.zeros(MLP_PARAMS, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    _check(library().artra_mlp_gradient(_fptr(p), _fptr(f), _fptr(y), len(y), tau, wpos, _fptr(g), _fptr(loss)))
    return g, float(loss[0])


def mlp_logits(params, features):
    p = np.ascontiguousarray(params, dtype=np.float32)
    f = np.ascontiguousarray(features, dtype=np.float32)
    out = np.zeros(len(f), dtype=np.float32)
    _check(library().artra_mlp_logits(_fptr(p), _fptr(f), len(f), _fptr(out)))
    return out


def calc(expr, x=0.0, y=0.0, z=0.0):
    """Evaluate an initial-condition expression; None where the solver's
    calculator rejects it (it then keeps its previous answer)."""
    out = ctypes.c_double(0.0)
    rc = library().artra_calc_expression(expr.encode(), x, y, z, ctypes.byref(out))
    return None if rc else out.value


def copy_probe(nbytes=1 << 30, iters=20):
    """Read+write GB/s of a plain device copy of nbytes bytes."""
    out = ctypes.c_double()
    _check(library().artra_hip_copy_probe(nbytes, iters, ctypes.byref(out)))
    return out.value


class Case:
    """A case directory prepared like the solver's Preprocess and
    InitializeComputeDomain (up to the first boundary treatment)."""

    def __init__(self, directory):
        lib = library()
        h = ctypes.c_void_p()
        _check(lib.artra_case_open(os.fspath(directory).encode(), ctypes.byref(h)))
        self._h = h
        self.directory = os.fspath(directory)
        n = (ctypes.c_int * 3)()
        ng = (ctypes.c_int * 3)()
        pin = (ctypes.c_int * 6)()
        lib.artra_case_grid(h, n, ng, pin)
        self.n = tuple(n)              # node counts (x, y, z) including ghosts
        self.ng = tuple(ng)
        self.pin = ((pin[0], pin[1]), (pin[2], pin[3]), (pin[4], pin[5]))

    @property
    def handle(self):
        return self._h

    @property
    def space(self):
        return library().artra_case_space(self._h)

    @property
    def model(self):
        return library().artra_case_model(self._h)

    @property
    def time(self):
        return library().artra_case_time(self._h)

    @property
    def nodes(self):
        return self.n[0] * self.n[1] * self.n[2]

    def field(self, level=TO):
        """Level of the host node array as [5][nz][ny][nx]."""
        out = np.empty(5 * self.nodes)
        library().artra_case_get_field(self._h, level, _dptr(out))
        return out.reshape(5, self.n[2], self.n[1], self.n[0])

    def set_field(self, level, field):
        arr = np.ascontiguousarray(field, dtype=np.float64).reshape(-1)
        if arr.size != 5 * self.nodes:
            raise ValueError("field has the wrong size")
        library().artra_case_set_field(self._h, level, _dptr(arr))

    def flags(self):
        """did, fid, lid, gst as [4][nz][ny][nx] int32."""
        out = np.empty(4 * self.nodes, dtype=np.int32)
        library().artra_case_get_flags(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out.reshape(4, self.n[2], self.n[1], self.n[0])

    def clock(self):
        """(now, steps taken) of the case's time loop."""
        now, step = ctypes.c_double(), ctypes.c_int()
        library().artra_case_clock(self._h, ctypes.byref(now), ctypes.byref(step))
        return now.value, step.value

    def forces(self):
        """Surface force of every body as [bodies][9] = Fp, Fv, Tt (the last
        integration's values)."""
        n = library().artra_case_forces(self._h, None)
        out = np.zeros((n, 9))
        library().artra_case_forces(self._h, _dptr(out))
        return out

    def bodies(self):
        """Motion state of every body as [bodies][15] = O, V, W, at, ar."""
        n = library().artra_case_bodies(self._h, None)
        out = np.zeros((n, 15))
        library().artra_case_bodies(self._h, _dptr(out))
        return out

    def interior(self, arr):
        """Restrict a [..., nz, ny, nx] array to the interior range."""
        (x0, x1), (y0, y1), (z0, z1) = self.pin
        return arr[..., z0:z1, y0:y1, x0:x1]

    def close(self):
        if self._h:
            library().artra_case_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """Device-resident solver attached to a Case (artra_hip_attach).

    ``initial_boundary`` applies TreatBoundary(TO) after the upload, which is
    the last step of the solver's InitializeComputeDomain."""

    def __init__(self, case, initial_boundary=True, segment=0, exact=None, gather=True, partial=True, fused=True,
                 graph=True):
        self.case = case
        lib = library()
        h = ctypes.c_void_p()
        _check(lib.artra_hip_attach(case.space, case.model, ctypes.byref(h)))
        self._h = h
        if exact is not None:
            _check(lib.artra_hip_set_arithmetic(h, 1 if exact else 0))
        if not (gather and partial):
            _check(lib.artra_hip_set_fill_options(h, 1 if gather else 0, 1 if partial else 0))
        if segment:
            _check(lib.artra_hip_set_tuning(h, segment))
        if not fused:
            _check(lib.artra_hip_set_fused(h, 0))
        if not graph:
            _check(lib.artra_hip_set_graph(h, 0))
        if initial_boundary:
            self.treat_boundary(TO)

    def treat_boundary(self, level):
        _check(library().artra_hip_treat_boundary(self._h, level))

    def compute_dt(self):
        dt = ctypes.c_double()
        _check(library().artra_hip_compute_dt(self._h, self.case.time, ctypes.byref(dt)))
        return dt.value

    def evolve(self, dt):
        _check(library().artra_hip_evolve(self._h, dt))

    def march(self, max_steps=2 ** 31 - 1):
        """The solver's time loop without output; returns (steps, dts)."""
        cap = min(max_steps, 1 << 20)
        dts = np.zeros(cap)
        done = ctypes.c_int(0)
        _check(library().artra_hip_march(self._h, self.case.time, cap, ctypes.byref(done), _dptr(dts)))
        return done.value, dts[: done.value].copy()

    def field(self, level=TO):
        out = np.empty(5 * self.case.nodes)
        _check(library().artra_hip_read_field(self._h, level, _dptr(out)))
        c = self.case
        return out.reshape(5, c.n[2], c.n[1], c.n[0])

    def set_field(self, level, field):
        arr = np.ascontiguousarray(field, dtype=np.float64).reshape(-1)
        _check(library().artra_hip_write_field(self._h, level, _dptr(arr)))

    def download(self, level=TO):
        _check(library().artra_hip_download(self._h, level))

    def upload(self, level=TO):
        _check(library().artra_hip_upload(self._h, level))

    def synchronize(self):
        _check(library().artra_hip_synchronize(self._h))

    def surface_force(self):
        """IntegrateSurfaceForce on the device; the case's bodies hold the result."""
        _check(library().artra_hip_surface_force(self._h))

    def write_probe_data(self, kind):
        """The probe files of one kind (0 point, 1 line, 2 curve, 3 surface
        force) in the current directory."""
        _check(library().artra_hip_write_probe_data(self._h, self.case.time, kind))

    def write_space_data(self):
        """The field, geometry and log files of the output in the current directory."""
        _check(library().artra_hip_write_space_data(self._h, self.case.time))

    def write_restart(self, path):
        """FP64 restart file of level TO and the clock."""
        _check(library().artra_hip_write_restart(self._h, self.case.time, os.fsencode(path)))

    def read_restart(self, path):
        """Continue from an FP64 restart file (field, clock, output recorder)."""
        _check(library().artra_hip_read_restart(self._h, self.case.time, os.fsencode(path)))

    def solve(self):
        """The solver loop with its output cadence, writing every output
        file into the current directory."""
        _check(library().artra_hip_solve(self._h, self.case.time))

    def set_timing(self, on=True):
        _check(library().artra_hip_set_timing(self._h, 1 if on else 0))

    def timing(self):
        """Per-category (sweep, boundary, immersed boundary, cfl): (ms, launches)."""
        ms = (ctypes.c_double * 4)()
        n = (ctypes.c_long * 4)()
        _check(library().artra_hip_get_timing(self._h, ms, n))
        names = ("sweep", "boundary", "immersed", "cfl")
        return {names[q]: (ms[q], n[q]) for q in range(4)}

    def fused_axes(self):
        """Axes (0 x, 1 y, 2 z) whose directional operator runs as one fused launch."""
        m = ctypes.c_int()
        _check(library().artra_hip_fused_axes(self._h, ctypes.byref(m)))
        return [p for p in range(3) if m.value >> p & 1]

    def sample_faces(self, axis):
        """Features [n][12] and labels [n] of every face along an axis (level TO)."""
        n = ctypes.c_long()
        _check(library().artra_hip_face_count(self._h, axis, ctypes.byref(n)))
        feat = np.zeros((n.value, 12), dtype=np.float32)
        lab = np.zeros(n.value, dtype=np.float32)
        _check(library().artra_hip_sample_faces(self._h, axis, _fptr(feat), _fptr(lab)))
        return feat, lab

    def set_indicator(self, params, logit_threshold=0.0, smooth_variation=1e-3):
        """Route WENO5 faces with the learned indicator (None: plain WENO5),
        classified once per directional operator; faces varying less than
        smooth_variation (relative) take the linear flux without the network
        (0: never)."""
        _check(library().artra_hip_set_indicator_options(self._h, 0, float(smooth_variation)))
        if params is None:
            _check(library().artra_hip_set_indicator(self._h, None, 0.0))
            return
        p = np.ascontiguousarray(params, dtype=np.float32)
        _check(library().artra_hip_set_indicator(self._h, _fptr(p), float(logit_threshold)))

    def hybrid_counts(self, reset=False):
        """(faces computed with the linear flux, faces classified, faces
        computed with the full WENO5 flux, faces the indicator flagged
        troubled); the routing is per wavefront of 64 faces."""
        c = (ctypes.c_ulonglong * 4)()
        _check(library().artra_hip_hybrid_counts(self._h, c, 1 if reset else 0))
        return tuple(int(v) for v in c)

    def close(self):
        if self._h:
            library().artra_hip_detach(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
