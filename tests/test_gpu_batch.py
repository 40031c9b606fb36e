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
    lin, seen, weno, _ = b.counts()
    assert 0 < lin < seen and lin + weno == seen
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


def _committed_indicator():
    from artracfd_amd import indicator
    params, meta = indicator.load()
    return params, float(meta.get("logit_threshold", indicator.LOGIT_THRESHOLD))


@pytest.mark.parametrize("T", [64, 128, 256])
def test_batch_with_trained_indicator_matches_single_runs_for_every_tile_shape(tmp_path, T):
    """The committed weights on oblique shocks: lines with and without
    troubled faces share tiles, each face is routed on its own, so every
    case matches its single-case run bit for bit whatever the tile shape (T
    threads per line, 256 / T lines per block)."""
    print(f"batch trained indicator T={T}", flush=True)
    from artracfd_amd import Case, Solver
    from artracfd_amd.batch import Batch
    params, thr = _committed_indicator()
    dirs = _cases(tmp_path, SPEC)
    ref = _single(dirs, 5, params, thr)
    cases = [Case(d) for d in dirs]
    solvers = [Solver(c) for c in cases]
    b = Batch(solvers)
    b.set_indicator(params, thr)
    b.set_threads([T] * len(dirs))
    b.counts(reset=True)
    b.march(5)
    lin, seen, weno, _ = b.counts()
    print("linear", lin / seen, "weno5", weno / seen)
    assert 0 < weno < seen and 0 < lin and lin + weno == seen
    for q, (c, s) in enumerate(zip(cases, solvers)):
        assert _now(c) == ref[q][1], q
        assert np.array_equal(c.interior(s.field()), ref[q][0]), q
    b.close()
    for s, c in zip(solvers, cases):
        s.close()
        c.close()


def test_batched_workload_seed2026_with_indicator(tmp_path):
    """The batched workload as benchmarked: 64 oblique-shock cases of sides
    32-128 drawn from seed 2026, tuned, 100 lockstep steps replayed from a
    HIP graph with the committed indicator.  Every case is bit-identical to
    its own single-case run, and every case's clock and step count match."""
    print("batched workload seed 2026, 64 cases x 100 steps, indicator on", flush=True)
    from artracfd_amd import Case, Solver
    from artracfd_amd.batch import Batch, oblique_shock_cases
    params, thr = _committed_indicator()
    dirs, drawn = oblique_shock_cases(str(tmp_path), 2026, 64, 100)
    assert sorted(set(d["side"] for d in drawn)) == [32, 48, 64, 96, 128]
    cases = [Case(d) for d in dirs]
    solvers = [Solver(c) for c in cases]
    b = Batch(solvers)
    b.set_indicator(params, thr)
    b.tune(reps=1)
    b.march(100, graph=True)
    got = [(c.interior(s.field()), c.clock()) for c, s in zip(cases, solvers)]
    b.close()
    for s, c in zip(solvers, cases):
        s.close()
        c.close()
    for q, d in enumerate(dirs):
        case = Case(d)
        solver = Solver(case)
        solver.set_indicator(params, thr)
        done, _ = solver.march(100)
        assert done == 100
        assert case.clock() == got[q][1], q
        assert np.array_equal(case.interior(solver.field()), got[q][0]), (q, drawn[q])
        solver.close()
        case.close()


def test_batched_smallest_cases_match_oracle(tmp_path, oracle_mod):
    """The two smallest cases of the seed-2026 draw, batched in plain WENO5
    for 5 steps, against the CPU oracle's march of the same cases (relative
    L2 per variable group <= 1e-10, dt to 1e-12)."""
    from conftest import rel_l2

    from artracfd_amd import Case, Solver
    from artracfd_amd.batch import Batch, draw, oblique_shock_case
    drawn = draw(2026, 64)
    small = sorted(range(64), key=lambda q: drawn[q]["side"])[:2]
    dirs = [oblique_shock_case(str(tmp_path / f"c{q}"), drawn[q]["side"], drawn[q]["mach"], drawn[q]["k"], steps=100)
            for q in small]
    print("batched vs oracle", [drawn[q] for q in small], flush=True)
    cases = [Case(d) for d in dirs]
    solvers = [Solver(c) for c in cases]
    b = Batch(solvers)
    b.march(5)
    got = [c.interior(s.field()) for c, s in zip(cases, solvers)]
    b.close()
    for s, c in zip(solvers, cases):
        s.close()
        c.close()
    for q, d in enumerate(dirs):
        case = Case(d)
        oracle_mod.treat_boundary(case, 0)
        n, _ = oracle_mod.march(case, 5)
        assert n == 5
        err = rel_l2(got[q], case.interior(case.field()))
        print(q, err)
        assert max(err.values()) <= 1e-10
        case.close()
