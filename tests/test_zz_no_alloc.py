"""The allocation-free march (artra_hip_march_noalloc) against artra_hip_march
through the C ABI: same step count, same dt sequence and the same field bit
for bit, on two calls in a row (the second on a smaller case, so the shared
per-GPU scratch holds the first call's contents), past the end time, and the
size limit refused on the host."""
import ctypes

import numpy as np
import pytest

from conftest import case_dir

pytestmark = pytest.mark.gpu


def _run(name, steps, noalloc):
    from artracfd_amd import Case, Solver, library
    case = Case(case_dir(name))
    solver = Solver(case)
    if noalloc:
        dts = np.zeros(steps)
        done = ctypes.c_int(-1)
        rc = library().artra_hip_march_noalloc(solver._h, case.time, steps, ctypes.byref(done),
                                               dts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert rc == 0, library().artra_hip_error()
        done, dts = done.value, dts[: done.value].copy()
    else:
        done, dts = solver.march(steps)
    U = case.interior(solver.field()).copy()
    clock = case.clock()
    solver.close()
    case.close()
    return done, dts, U, clock


@pytest.mark.parametrize("name,steps", [("shockvortex64", 12), ("sod", 8), ("sod", 200)])
def test_noalloc_march_matches_march(name, steps):
    """shockvortex64 then the smaller Sod tube (the scratch keeps the larger
    case's contents between the calls); Sod for 200 steps reaches its end
    time after 186, and the steps past it change nothing."""
    print(f"no-alloc march: {name}, {steps} steps", flush=True)
    d0, dt0, U0, c0 = _run(name, steps, noalloc=False)
    d1, dt1, U1, c1 = _run(name, steps, noalloc=True)
    assert d1 == d0
    assert np.array_equal(dt1, dt0)
    assert np.array_equal(U1, U0)
    assert c1 == c0
    if name == "sod" and steps == 200:
        assert d0 == 186


def test_noalloc_march_refuses_too_many_steps():
    from artracfd_amd import Case, Solver, library
    case = Case(case_dir("sod"))
    solver = Solver(case)
    lib = library()
    done = ctypes.c_int(-1)
    limit = 1024
    rc = lib.artra_hip_march_noalloc(solver._h, case.time, limit + 1, ctypes.byref(done), None)
    assert rc != 0 and b"max_steps" in lib.artra_hip_error()
    assert done.value == 0
    assert case.clock() == (0.0, 0)
    solver.close()
    case.close()
