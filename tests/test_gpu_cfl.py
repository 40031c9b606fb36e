"""The next step's CFL maxima taken from the epilogue of the step's last
fused operator: the same dt as the separate reduction over the field, and
not used once the field has been written in between."""
import numpy as np
import pytest

from conftest import case_dir

pytestmark = pytest.mark.gpu


def _solver(name):
    from artracfd_amd import Case, Solver
    case = Case(case_dir(name))
    return case, Solver(case)


@pytest.mark.parametrize("name", ["shockvortex64", "sod"])
def test_fused_cfl_matches_the_field_reduction(name):
    print("fused cfl", name, flush=True)
    case, s = _solver(name)
    s.set_timing(True)
    s.march(3)
    assert s.fused_axes() and s.timing()["cfl"][1] == 1   # the separate reduction ran for the first step only
    s.set_timing(False)
    dt_fused = s.compute_dt()           # from the last operator's epilogue
    dt_field = s.compute_dt()           # the slots are used once: the separate reduction
    assert dt_fused == dt_field
    s.close()
    case.close()


def test_field_written_between_steps_is_used():
    from artracfd_amd.lib import TO
    print("fused cfl after a field write", flush=True)
    case, s = _solver("shockvortex64")
    s.march(2)
    U = s.field()
    U[1:4] *= 1.5                       # faster flow: a smaller dt
    s.set_field(TO, U)
    dt_written = s.compute_dt()
    dt_again = s.compute_dt()
    assert dt_written == dt_again
    s.march(1)
    s.set_field(TO, s.field())          # same values, slots dropped
    assert s.compute_dt() == s.compute_dt()
    s.close()
    case.close()
