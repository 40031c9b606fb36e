"""Batched cases: many independent cases advanced in lockstep, one
persistent launch per directional operator over the lines of every case
(artra_hip_batch_*), optionally replayed from a captured HIP graph.

The synthetic workload of the batched benchmark: oblique shocks running
into randomly perturbed gas, cubic grids of uneven sizes, all drawn from one
seed (oblique_shock_cases).
"""
import ctypes
import os

import numpy as np

from . import lib
from .lib import ArtraError, Case, Solver, _check, _fptr

SIDES = (32, 48, 64, 96, 128)
GAMMA = 1.4


def _bind():
    L = lib.library()
    if getattr(L, "_batch_bound", False):
        return L
    vp = ctypes.c_void_p
    ip = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)
    sig = {
        "artra_hip_batch_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.c_int,
                                                  ctypes.POINTER(vp)]),
        "artra_hip_batch_destroy": (ctypes.c_int, [vp]),
        "artra_hip_batch_tune": (ctypes.c_int, [vp, ctypes.c_int, ip, ip]),
        "artra_hip_batch_set_threads": (ctypes.c_int, [vp, ip]),
        "artra_hip_batch_set_indicator": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_float]),
        "artra_hip_batch_counts": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_int]),
        "artra_hip_batch_march": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    L._batch_bound = True
    return L


class Batch:
    """Lockstep batch over already attached solvers (one per case)."""

    def __init__(self, solvers):
        L = _bind()
        self.solvers = list(solvers)
        n = len(self.solvers)
        sv = (ctypes.c_void_p * n)(*[s._h.value for s in self.solvers])
        tv = (ctypes.c_void_p * n)(*[s.case.time for s in self.solvers])
        h = ctypes.c_void_p()
        _check(L.artra_hip_batch_create(sv, tv, n, ctypes.byref(h)))
        self._h = h

    def tune(self, reps=2):
        """Autotune tile shapes and the persistent grid; returns (threads per line per case, grid)."""
        n = len(self.solvers)
        threads = (ctypes.c_int * n)()
        grid = ctypes.c_int()
        _check(_bind().artra_hip_batch_tune(self._h, reps, threads, ctypes.byref(grid)))
        return list(threads), grid.value

    def set_threads(self, threads):
        """Tile shape per case (threads per line) instead of the tuner's choice."""
        n = len(self.solvers)
        t = (ctypes.c_int * n)(*[int(v) for v in threads])
        _check(_bind().artra_hip_batch_set_threads(self._h, t))

    def set_indicator(self, params, logit_threshold=0.0):
        if params is None:
            _check(_bind().artra_hip_batch_set_indicator(self._h, None, 0.0))
            return
        p = np.ascontiguousarray(params, dtype=np.float32)
        _check(_bind().artra_hip_batch_set_indicator(self._h, _fptr(p), float(logit_threshold)))

    def counts(self, reset=False):
        """(faces on the linear flux, faces classified, faces on the full
        WENO5 flux, faces flagged troubled), as Solver.hybrid_counts."""
        c = (ctypes.c_ulonglong * 4)()
        _check(_bind().artra_hip_batch_counts(self._h, c, 1 if reset else 0))
        return tuple(int(v) for v in c)

    def march(self, steps, graph=True):
        """steps lockstep time steps of every case; returns the device milliseconds."""
        ms = ctypes.c_double()
        _check(_bind().artra_hip_batch_march(self._h, int(steps), 1 if graph else 0, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if self._h:
            _bind().artra_hip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def post_shock(ms, rho1=1.0, p1=1.0):
    """Rankine-Hugoniot state behind a shock of Mach ms running into gas at rest."""
    g = GAMMA
    rho2 = rho1 * (g + 1) * ms ** 2 / ((g - 1) * ms ** 2 + 2)
    p2 = p1 * (1 + 2 * g / (g + 1) * (ms ** 2 - 1))
    u2 = ms * np.sqrt(g * p1 / rho1) * (1 - rho1 / rho2)
    return rho2, u2, p2


def draw(seed=2026, count=64):
    """Case parameters in a fixed draw order: side, shock Mach, wavenumbers."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        side = int(rng.choice(SIDES))
        ms = float(rng.uniform(1.2, 3.0))
        k = [int(v) for v in rng.integers(1, 9, 3)]
        out.append({"side": side, "mach": ms, "k": k})
    return out


def oblique_shock_case(directory, side, mach, k, steps=100, angle_deg=30.0):
    """Unit cube, outflow on every face.  Gas at rest with p = 1 and density
    1 + 0.1 sin(2 pi (kx x + ky y + kz z)); a plane shock of Mach `mach`
    through (0.3, 0.5, 0.5) whose normal lies in the x-y plane at angle_deg
    to x, with the Rankine-Hugoniot state behind it (pre-shock density 1)."""
    os.makedirs(directory, exist_ok=True)
    rho2, u2, p2 = post_shock(mach)
    a = np.radians(angle_deg)
    nx, ny = np.cos(a), np.sin(a)
    lines = ["space begin", "0, 0, 0", "1, 1, 1", f"{side}, {side}, {side}", "space end",
             "time begin", "0", "1000000000.0", "0.6", str(steps), "1", "0", "time end",
             "numerical begin", "1", "1", "0", "0", "0", "0", "1", "numerical end",
             "material begin", "0", "0.0", "0", "0, 0, 0", "material end",
             "reference begin", "1.0", "1.0", "1.0", "1.0", "reference end",
             "initialization begin", f"1+0.1*sin(2*pi*({k[0]}*x+{k[1]}*y+{k[2]}*z))", "0", "0", "0", "1",
             "initialization end"]
    for name in ("west", "east", "south", "north", "front", "back"):
        lines += [f"{name} boundary begin", "outflow", f"{name} boundary end"]
    lines += ["plane initialization begin", "0.3, 0.5, 0.5", f"{-nx:.9f}, {-ny:.9f}, 0",
              f"{rho2:.9f}", f"{u2 * nx:.9f}", f"{u2 * ny:.9f}", "0", f"{p2:.9f}", "plane initialization end"]
    with open(os.path.join(directory, "artracfd.case"), "w") as fp:
        fp.write("\n".join(lines) + "\n")
    with open(os.path.join(directory, "artracfd.geo"), "w") as fp:
        fp.write("count begin\n0\n0\ncount end\n")
    return directory


def oblique_shock_cases(root, seed=2026, count=64, steps=100):
    """Write the batched workload under root; returns (directories, parameters)."""
    params = draw(seed, count)
    dirs = [oblique_shock_case(os.path.join(root, f"case{c:03d}"), steps=steps, **p) for c, p in enumerate(params)]
    return dirs, params
