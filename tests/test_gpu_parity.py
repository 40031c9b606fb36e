"""Parity of the HIP path with the compiled reference (golden fixtures) and
with the CPU oracle on the same inputs.

The contract is 1e-10 relative L2 per variable group (rho, momentum,
energy) after the stated step counts.  The device kernels keep the
reference's operation order under -ffp-contract=off, so every node is also
checked to agree within 1e-12 of its component's magnitude, and the time
steps within 1e-13 relative.  Inviscid runs are expected bit-identical; the
viscous flux calls pow(), whose device and glibc results may differ in the
last bit, so viscous runs agree to rounding level instead."""
import ctypes
import os

import numpy as np
import pytest

from conftest import GOLDEN, case_dir, golden, rel_l2

pytestmark = pytest.mark.gpu

SMALL = ["sod", "tgv16", "cyl64", "sphere20", "v_base", "v_weno3", "v_rk2", "v_roe", "v_sw", "v_unsplit",
         "v_mix", "v_visc", "v_collapse", "v_walls", "v_gravity", "v_gravity2"]


def assert_node_agreement(a, b, tol=1e-12):
    """Every node within tol of its variable group's largest magnitude
    (density, momentum vector, energy)."""
    for sl in (slice(0, 1), slice(1, 4), slice(4, 5)):
        scale = max(float(np.abs(b[sl]).max()), 1e-300)
        diff = float(np.abs(a[sl] - b[sl]).max())
        assert diff <= tol * scale, (sl, diff, scale)


def _open(name, exact=None):
    from artracfd_amd import Case, Solver
    case = Case(case_dir(name))
    return case, Solver(case, exact=exact)


@pytest.mark.parametrize("name", SMALL)
@pytest.mark.parametrize("arith", ["exact", "fast"])
def test_case_matches_reference(name, arith):
    """exact: the reference's operation order; fast: the default fused form."""
    print(f"case {name} {arith}", flush=True)
    g = golden(name)
    case, solver = _open(name, exact=(arith == "exact"))
    U0 = case.interior(solver.field())
    assert np.array_equal(U0, g["U_init"]), "initial field after TreatBoundary(TO) differs"
    steps = int(g["steps"])
    done, dts = solver.march(steps)
    assert done == steps
    np.testing.assert_allclose(dts, g["dt"], rtol=1e-13 if arith == "exact" else 1e-11, atol=0)
    U = case.interior(solver.field())
    err = rel_l2(U, g["U_final"])
    print(name, arith, err, "bitwise:", np.array_equal(U, g["U_final"]), "dt bitwise:", np.array_equal(dts, g["dt"]))
    assert max(err.values()) <= 1e-10
    assert_node_agreement(U, g["U_final"], 1e-12 if arith == "exact" else 1e-10)
    solver.close()
    case.close()


@pytest.mark.parametrize("name", ["v_walls", "cyl64", "sphere20", "v_collapse", "sod", "tgv16"])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_treat_boundary_matches_oracle(name, level, oracle_mod):
    """Random interior states on every level; TreatBoundary on the device and
    in the oracle must give the same ghost values."""
    print(f"boundary {name} level {level}", flush=True)
    from artracfd_amd import Case, Solver
    case = Case(case_dir(name))
    rng = np.random.default_rng(7 + level)
    base = case.field(0)
    fl = case.flags()
    U = base.copy()
    noise = 1.0 + 0.05 * rng.standard_normal(U.shape)
    inner = case.interior(fl[0]) >= 0
    view = case.interior(U)
    view[:, inner] = view[:, inner] * case.interior(noise)[:, inner]
    case.set_field(level, U)
    solver = Solver(case, initial_boundary=False)
    solver.set_field(level, U)
    solver.treat_boundary(level)
    dev = solver.field(level)
    oracle_mod.treat_boundary(case, level)
    ref = case.field(level)
    np.testing.assert_array_equal(dev, ref)
    solver.close()
    case.close()


def test_dropin_evolve_matches_oracle(oracle_mod):
    """The drop-in EvolveFluidDynamics(dt, Space*, Model*) on the host node
    array against the oracle's restatement on an identical copy."""
    from artracfd_amd import Case, library
    a, b = Case(case_dir("v_visc")), Case(case_dir("v_visc"))
    oracle_mod.treat_boundary(a, 0)
    oracle_mod.treat_boundary(b, 0)
    lib = library()
    for _ in range(3):
        dt = oracle_mod.compute_dt(a)
        lib.EvolveFluidDynamics(dt, a.space, a.model)
        oracle_mod.evolve(b, dt)
    assert_node_agreement(a.field(0), b.field(0))
    a.close()
    b.close()


@pytest.mark.parametrize("name, level", [("v_walls", 1), ("cyl64", 0), ("sphere20", 2)])
def test_dropin_treat_boundary_matches_oracle(name, level, oracle_mod):
    """The drop-in TreatBoundary(tn, Space*, Model*) symbol on the host node
    array (walls, the immersed-boundary fill of a triangulated and of an
    analytic body) against the oracle's restatement on an identical copy
    with the same perturbed interior: bit for bit."""
    print(f"drop-in TreatBoundary {name} level {level}", flush=True)
    from artracfd_amd import Case, library
    a, b = Case(case_dir(name)), Case(case_dir(name))
    rng = np.random.default_rng(3 + level)
    U = a.field(0)
    fl = a.flags()
    inner = a.interior(fl[0]) >= 0
    noise = 1.0 + 0.05 * rng.standard_normal(U.shape)
    view = a.interior(U)
    view[:, inner] = view[:, inner] * a.interior(noise)[:, inner]
    a.set_field(level, U)
    b.set_field(level, U)
    library().TreatBoundary(level, a.space, a.model)
    oracle_mod.treat_boundary(b, level)
    np.testing.assert_array_equal(a.field(level), b.field(level))
    a.close()
    b.close()


DROPIN = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "artra_dropin")


@pytest.mark.skipif(not os.path.exists(DROPIN), reason="the drop-in driver is built only where the reference is")
@pytest.mark.parametrize("name", ["cyl64", "v_walls", "v_gravity"])
def test_reference_driver_through_dropin_symbols(name, tmp_path, oracle_mod):
    """The reference's own driver - Preprocess, InitializeComputeDomain, its
    ComputeTimeStep and time loop - linked by the INTEGRATION.md recipe, so
    that its ComputeGeometricField, TreatBoundary and EvolveFluidDynamics
    calls run in libartra_hip (the field on the GPU).  With the reference's
    operation order (ARTRA_EXACT=1) the run equals the reference's CPU run
    bit for bit: node flags, every dt and the final field."""
    import shutil
    import subprocess
    print(f"reference driver through the drop-in symbols: {name}", flush=True)
    g = golden(name)
    work = tmp_path / name
    shutil.copytree(case_dir(name), work)
    env = dict(os.environ, ARTRA_EXACT="1")
    out = subprocess.run([DROPIN, "run", str(int(g["steps"])), "out"], cwd=work, env=env, capture_output=True,
                         text=True, timeout=600)
    assert out.returncode == 0, out.stderr
    _, flags, Ui = oracle_mod.read_dump(str(work / "out_init.bin"))
    hdr, _, Uf = oracle_mod.read_dump(str(work / "out_final.bin"))
    assert hdr["steps"] == int(g["steps"])
    assert np.array_equal(flags.astype(np.int16), g["flags"])
    np.testing.assert_array_equal(np.fromfile(str(work / "out_dt.bin")), g["dt"])
    (x0, x1), (y0, y1), (z0, z1) = g["pin"]
    assert np.array_equal(Ui[:, z0:z1, y0:y1, x0:x1], g["U_init"])
    assert np.array_equal(Uf[:, z0:z1, y0:y1, x0:x1], g["U_final"])


@pytest.mark.skipif(not os.path.exists(DROPIN), reason="the drop-in driver is built only where the reference is")
@pytest.mark.parametrize("name", ["spheres_collide", "ico_moving"])
def test_reference_solve_with_moving_bodies_through_dropin_symbols(name, tmp_path, oracle_mod):
    """The reference's whole Solve with moving bodies (two colliding spheres;
    a tumbling triangulated sphere), its own solid
    dynamics (surface force, kinematics, collisions, motion) calling the
    library's ComputeGeometricField (classification and reconstruction of
    uncovered nodes on the caller's node array) and TreatImmersedBoundary,
    and its time loop calling EvolveFluidDynamics on the GPU: in exact
    arithmetic the final field, surface forces and body states are the
    reference's CPU run's bit for bit."""
    import shutil
    import subprocess
    print(f"reference solve through the drop-in symbols: {name}", flush=True)
    fx = golden("solve_" + name)
    work = tmp_path / name
    shutil.copytree(case_dir(name), work)
    env = dict(os.environ, ARTRA_EXACT="1")
    out = subprocess.run([DROPIN, "solve", "out"], cwd=work, env=env, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr
    hdr, _, Uf = oracle_mod.read_dump(str(work / "out_final.bin"))
    assert hdr["steps"] == int(fx["steps"])
    from artracfd_amd import Case
    case = Case(case_dir(name))
    U = case.interior(Uf)
    case.close()
    assert np.array_equal(U, fx["U_final"])
    assert np.array_equal(np.fromfile(str(work / "out_force.bin")).reshape(-1, 9), fx["force"])
    assert np.array_equal(np.fromfile(str(work / "out_bodies.bin")).reshape(-1, 15), fx["bodies"])


def test_compute_dt_matches_oracle(oracle_mod):
    from artracfd_amd import Case, Solver
    case = Case(case_dir("cyl64"))
    solver = Solver(case)
    solver.download(0)
    assert solver.compute_dt() == oracle_mod.compute_dt(case)
    solver.close()
    case.close()


@pytest.mark.parametrize("name", ["sod", "tgv16", "cyl64", "v_walls", "v_rk2", "v_unsplit"])
def test_fill_strategies_give_identical_results(name):
    """The resolved gather and the axis-only fills between stages must give
    the same field as the literal ordered fill after every stage."""
    print(f"fill strategies {name}", flush=True)
    from artracfd_amd import Case, Solver
    fields = []
    for gather, partial in ((False, False), (True, True), (True, False), (False, True)):
        case = Case(case_dir(name))
        solver = Solver(case, exact=True, gather=gather, partial=partial)
        solver.march(6)
        fields.append(solver.field())
        solver.close()
        case.close()
    for f in fields[1:]:
        np.testing.assert_array_equal(case_interior(name, f), case_interior(name, fields[0]))


@pytest.mark.parametrize("name", ["sod", "tgv16", "cyl64", "v_gravity", "v_walls", "shockvortex64"])
def test_graph_replayed_steps_equal_launched_steps(name):
    """The device march replaying each step from a captured graph gives the
    field and dt sequence of the march that launches the kernels one by one,
    across march calls and after a change of arithmetic between them (the
    graph is captured again)."""
    print(f"graph replay {name}", flush=True)
    from artracfd_amd import Case, Solver, library
    runs = []
    for graph in (True, False):
        case = Case(case_dir(name))
        solver = Solver(case, graph=graph)
        _, d1 = solver.march(4)
        _, d2 = solver.march(3)
        library().artra_hip_set_arithmetic(solver._h, 1)
        _, d3 = solver.march(3)
        runs.append((np.concatenate([d1, d2, d3]), case.interior(solver.field()).copy()))
        solver.close()
        case.close()
    assert len(runs[0][0]) == 10 and np.array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])


def case_interior(name, arr):
    g = golden(name)
    (x0, x1), (y0, y1), (z0, z1) = g["pin"]
    return arr[..., z0:z1, y0:y1, x0:x1]


def _variant(tmp_path, name, numerical):
    """Copy of a case directory with its numerical section replaced."""
    import shutil
    d = tmp_path / f"{name}_{''.join(map(str, numerical))}"
    shutil.copytree(case_dir(name), d)
    text = open(d / "artracfd.case").read()
    head, rest = text.split("numerical begin\n", 1)
    _, tail = rest.split("numerical end\n", 1)
    body = "".join(f"{v}\n" for v in numerical)
    open(d / "artracfd.case", "w").write(head + "numerical begin\n" + body + "numerical end\n" + tail)
    return str(d)


FUSED = [("sod", None), ("shockvortex64", None), ("sod", (0, 1, 0, 0, 0, 0, 1)), ("sod", (1, 0, 0, 0, 0, 0, 1)),
         ("sod", (1, 1, 0, 1, 0, 0, 1)), ("sod", (1, 1, 0, 0, 1, 0, 1)), ("shockvortex64", (0, 0, 0, 1, 1, 0, 1))]


@pytest.mark.parametrize("name, numerical", FUSED)
@pytest.mark.parametrize("arith", ["exact", "fast"])
def test_fused_line_sweeps_match_stage_by_stage_sweeps(tmp_path, name, numerical, arith):
    """One launch per directional operator with the line in LDS (rk_line)
    against the stage-by-stage sweeps with their between-stage fills: the
    final level must agree bit for bit (RK3/RK2, WENO5/WENO3, mean/Roe,
    LLF/Steger-Warming)."""
    print(f"fused {name} {numerical} {arith}", flush=True)
    from artracfd_amd import Case, Solver
    d = case_dir(name) if numerical is None else _variant(tmp_path, name, numerical)
    out = []
    for fused in (True, False):
        case = Case(d)
        solver = Solver(case, exact=(arith == "exact"), fused=fused)
        axes = solver.fused_axes()
        if fused:
            assert axes == [0, 1, 2], axes
        else:
            assert axes == []
        done, dts = solver.march(5)
        out.append((case.interior(solver.field()), dts))
        solver.close()
        case.close()
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][0], out[1][0])


FULLSIZE = [n for n in ("tgv128", "cylinder") if os.path.exists(os.path.join(GOLDEN, f"{n}_full.npz"))]


@pytest.mark.parametrize("name", FULLSIZE)
def test_fullsize_run_matches_reference(name):
    """The full-size runs (TGV 128^3 viscous, 200 steps; shock-cylinder
    512x512x4 with the immersed boundary, 1000 steps) in the default fast
    arithmetic against fixtures of the compiled reference's own run: the dt
    sequence, a strided subsample of the field (relative L2 per variable
    group <= 1e-10, the parity contract) and, over every interior fluid
    node, per-component sums and sums of squares within 1e-10 of their
    variable group's scale (sqrt(N * sum of squares) for a sum, the group's
    sum of squares for a sum of squares), which a relative L2 error of 1e-10
    respects."""
    print(f"full size {name}", flush=True)
    from artracfd_amd import Case, Solver
    g = golden(name + "_full")
    case = Case(case_dir(name))
    solver = Solver(case)
    done, dts = solver.march(int(g["steps"]))
    assert done == int(g["steps"])
    np.testing.assert_allclose(dts, g["dt"], rtol=1e-11, atol=0)
    U = case.interior(solver.field())
    fluid = case.interior(case.flags()[0]) == 0
    n = int(fluid.sum())
    for grp in ((0,), (1, 2, 3), (4,)):
        sq = sum(float(g["squares"][c]) for c in grp)
        for c in grp:
            s, q = U[c][fluid].sum(), (U[c][fluid] ** 2).sum()
            assert abs(s - g["sums"][c]) <= 1e-10 * np.sqrt(n * sq), (c, s, g["sums"][c])
            assert abs(q - g["squares"][c]) <= 1e-10 * sq, (c, q, g["squares"][c])
    st = int(g["stride"])
    sub = U[:, ::st, ::st, ::st]
    err = rel_l2(sub, g["sub"])
    print(name, "subsample rel L2", err)
    assert max(err.values()) <= 1e-10
    solver.close()
    case.close()


@pytest.mark.skipif("cylinder" not in FULLSIZE, reason="no full-size cylinder fixture")
def test_fullsize_cylinder_exact_mode_is_bitwise():
    """The inviscid full-size shock-cylinder run (512x512x4, 1000 steps,
    immersed boundary) in the reference's operation order: every dt and the
    whole interior field (SHA-256 of its FP64 bytes) equal the compiled
    reference's run bit for bit."""
    import hashlib
    print("full size cylinder, exact arithmetic, bitwise", flush=True)
    from artracfd_amd import Case, Solver
    g = golden("cylinder_full")
    case = Case(case_dir("cylinder"))
    solver = Solver(case, exact=True)
    done, dts = solver.march(int(g["steps"]))
    assert done == int(g["steps"])
    U = np.ascontiguousarray(case.interior(solver.field()))
    solver.close()
    case.close()
    assert np.array_equal(dts, g["dt"])
    assert hashlib.sha256(U.tobytes()).hexdigest() == str(g["sha256"])
