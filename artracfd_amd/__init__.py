"""MI355X implementation of ArtraCFD's fluid-update hot path.

The compute path is libartra_hip.so (HIP kernels for gfx950 behind a C ABI,
see include/artra_hip.h).  This package is a thin ctypes layer over that
library for tests, the benchmark and scripting:

    from artracfd_amd import Case, Solver
    case = Case("cases/sod")            # Preprocess + initialisation on the host
    solver = Solver(case)               # attach: field to HBM, TreatBoundary(TO)
    steps, dts = solver.march()         # EvolveSolution's loop on the device
    rho = solver.field()[0]             # level TO, component-major, [5][nz][ny][nx]

Importing the package never falls back to another implementation: if the
library is missing, loading it raises.
"""
from .lib import ArtraError, Case, Solver, library, library_path  # noqa: F401

__all__ = ["ArtraError", "Case", "Solver", "library", "library_path"]
