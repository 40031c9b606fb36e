"""Batched cases: one persistent launch per directional operator over the
lines of many independent cases must give, case by case, exactly what the
single-case fused path gives (same dt sequence, same fields)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cases(tmp_path, spec):
    from artracfd_amd import batch
    return [batch.oblique_shock_case(str(tmp_path / f"c{q}"), side, mach, k, steps=20)
            for q, (side, mach, k) in enumerate(spec)]


SPEC = [(32, 2.4, [1, 2, 3]), (48, 1.3, [8, 1, 4]), (32, 3.0, [5, 5, 2]), (64, 1.9, [2, 7, 1])]


def _single(dirs, steps, params=None, threshold=0.0):
    from artracfd_amd import Case, Solver
    out = []
    for d in dirs:
        case = Case(d)
        solver = Solver(case)
        if params is not None:
            solver.set_indicator(params, threshold)
        solver.march(steps)
        out.append((case.interior(solver.field()), case.clock()[0]))
        solver.close()
        case.close()
    return out


def _now(case):
    return case.clock()[0]


@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_single_case_runs(tmp_path, graph):
    print(f"batch vs single graph={graph}", flush=True)
    from artracfd_amd import Case, Solver
    from artracfd_amd.batch import Batch
    dirs = _cases(tmp_path, SPEC)
    ref = _single(dirs, 5)
    cases = [Case(d) for d in dirs]
    solvers = [Solver(c) for c in cases]
    b = Batch(solvers)
    b.march(3, graph=graph)
    b.march(2, graph=graph)
    for q, (c, s) in enumerate(zip(cases, solvers)):
        U = c.interior(s.field())
        assert _now(c) == ref[q][1], q
        assert np.array_equal(U, ref[q][0]), q
    b.close()
    for s, c in zip(solvers, cases):
        s.close()
        c.close()


def test_batch_tuner_keeps_the_field_and_the_result(tmp_path):
    print("batch tuner", flush=True)
    from artracfd_amd import Case, Solver
    from artracfd_amd.batch import Batch
    dirs = _cases(tmp_path, SPEC)
    ref = _single(dirs, 4)
    cases = [Case(d) for d in dirs]
    solvers = [Solver(c) for c in cases]
    before = [c.interior(s.field()) for c, s in zip(cases, solvers)]
    b = Batch(solvers)
    threads, grid = b.tune(reps=1)
    assert len(threads) == len(dirs) and all(t in (64, 128, 256) for t in threads) and grid > 0
    for q, (c, s) in enumerate(zip(cases, solvers)):
        assert np.array_equal(c.interior(s.field()), before[q])
    b.march(4)
    for q, (c, s) in enumerate(zip(cases, solvers)):
        assert np.array_equal(c.interior(s.field()), ref[q][0]), q
    b.close()
    for s, c in zip(solvers, cases):
        s.close()
        c.close()


def test_batch_with_indicator_matches_single_case_runs(tmp_path):
    print("batch hybrid vs single hybrid", flush=True)
    import mlp_oracle as M

    from artracfd_amd import Case, Solver
    from artracfd_amd.batch import Batch
    params = M.init_params(11)
    dirs = _cases(tmp_path, SPEC)
    ref = _single(dirs, 3, params, 0.5)
    cases = [Case(d) for d in dirs]
    solvers = [Solver(c) for c in cases]
    b = Batch(solvers)
    b.set_indicator(params, 0.5)
    b.march(3)
    lin, seen = b.counts()
    assert 0 < lin < seen
    for q, (c, s) in enumerate(zip(cases, solvers)):
        assert np.array_equal(c.interior(s.field()), ref[q][0]), q
    b.close()
    for s, c in zip(solvers, cases):
        s.close()
        c.close()


def test_batch_rejects_unsupported_cases():
    from conftest import case_dir

    from artracfd_amd import ArtraError, Case, Solver
    from artracfd_amd.batch import Batch
    case = Case(case_dir("tgv16"))            # viscous: no fused sweeps
    solver = Solver(case)
    with pytest.raises(ArtraError):
        Batch([solver])
    solver.close()
    case.close()
