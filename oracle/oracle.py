"""ctypes binding of liboracle.so, the CPU restatement of the hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.  The restatement works on
the solver's own node array (the Space/Model/Time of a prepared case), so a
case opened with artracfd_amd.Case can be advanced by either side.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF = os.path.join(HERE, "_ref", "artra_ref")
TO, TN, TM = 0, 1, 2

_lib = None


def build():
    """Compile liboracle.so, and where the reference sources exist the
    reference harness and the drop-in driver (the reference's objects linked
    with libartra_hip.so, which must be built first)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref", "-j8"], check=True)
        subprocess.run(["make", "-s", "-C", HERE, "dropin"], check=True)


def library():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
        lib.oracle_weno5.argtypes = [dp, dp]
        lib.oracle_weno3.argtypes = [dp, dp]
        lib.oracle_compute_fhat.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_int), vp, vp, dp]
        lib.oracle_compute_fvhat.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_int), dp, vp, vp, dp]
        lib.oracle_treat_boundary.argtypes = [ctypes.c_int, vp, vp]
        lib.oracle_evolve_fluid_dynamics.argtypes = [ctypes.c_double, vp, vp]
        lib.oracle_compute_time_step.argtypes = [vp, vp, vp]
        lib.oracle_compute_time_step.restype = ctypes.c_double
        lib.oracle_march.argtypes = [vp, vp, vp, ctypes.c_int, dp]
        lib.oracle_march.restype = ctypes.c_int
        lib.oracle_map_primitive.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp]
        lib.oracle_map_conservative.argtypes = [ctypes.c_double, dp, dp]
        lib.oracle_integrate_surface_force.argtypes = [vp, vp]
        lib.oracle_point_triangle_distance.argtypes = [dp] * 5
        lib.oracle_point_triangle_distance.restype = ctypes.c_double
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def weno5(F):
    """F: [..., 5, 5] (stencil point, component) -> [..., 5]."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    out = np.empty(F.shape[:-2] + (5,))
    flat_in, flat_out = F.reshape(-1, 25), out.reshape(-1, 5)
    for q in range(flat_in.shape[0]):
        library().oracle_weno5(_dp(flat_in[q]), _dp(flat_out[q]))
    return out


def weno3(F):
    F = np.ascontiguousarray(F, dtype=np.float64)
    out = np.empty(F.shape[:-2] + (5,))
    flat_in, flat_out = F.reshape(-1, 15), out.reshape(-1, 5)
    for q in range(flat_in.shape[0]):
        library().oracle_weno3(_dp(flat_in[q]), _dp(flat_out[q]))
    return out


def treat_boundary(case, level):
    library().oracle_treat_boundary(level, case.space, case.model)


def evolve(case, dt):
    library().oracle_evolve_fluid_dynamics(dt, case.space, case.model)


def compute_dt(case):
    return library().oracle_compute_time_step(case.time, case.space, case.model)


def surface_force(case):
    """IntegrateSurfaceForce on the case's level TO; returns [bodies][9]
    (Fp, Fv, Tt) as the case's bodies then hold them."""
    library().oracle_integrate_surface_force(case.space, case.model)
    return case.forces()


def march(case, max_steps=2 ** 31 - 1):
    """The solver's time loop on the case's node array; returns (steps, dts)."""
    cap = min(max_steps, 1 << 20)
    dts = np.zeros(cap)
    n = library().oracle_march(case.time, case.space, case.model, cap, _dp(dts))
    return n, dts[:n].copy()


def run_reference(case_dir, workdir, max_steps=0):
    """Run the compiled reference harness on a copy of case_dir (only where
    oracle/_ref was built).  Returns the paths of its dumps."""
    import shutil
    import json
    if os.path.exists(workdir):
        shutil.rmtree(workdir)
    shutil.copytree(case_dir, workdir)
    out = subprocess.run([REF, "run", str(max_steps), "out"], cwd=workdir, check=True, capture_output=True, text=True)
    info = json.loads(out.stdout.strip().splitlines()[-1])
    return info, {k: os.path.join(workdir, f"out_{k}.bin") for k in ("init", "final", "dt")}


def read_dump(path):
    """Read an ARTRAST1 dump: (header dict, flags[4][nz][ny][nx], U[5][nz][ny][nx])."""
    with open(path, "rb") as fp:
        magic = fp.read(8)
        if magic != b"ARTRAST1":
            raise ValueError("not an ARTRAST1 dump")
        hdr = np.frombuffer(fp.read(32), dtype=np.int32)
        t = np.frombuffer(fp.read(16), dtype=np.float64)
        n = (int(hdr[0]), int(hdr[1]), int(hdr[2]))
        N = n[0] * n[1] * n[2]
        flags = np.frombuffer(fp.read(16 * N), dtype=np.int32).reshape(4, n[2], n[1], n[0])
        U = np.frombuffer(fp.read(40 * N), dtype=np.float64).reshape(5, n[2], n[1], n[0])
    return {"n": n, "ng": tuple(int(v) for v in hdr[3:6]), "steps": int(hdr[6]), "now": float(t[0])}, flags, U
