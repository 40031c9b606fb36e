"""The solver loop on the device with its output cadence (artra_hip_solve):
every probe, surface-force, field, geometry and log file it writes against
the files the compiled reference's own Solve wrote for the same case
(fixtures tests/golden/solve_*.npz), the surface force itself
(solid_dynamics.c:50-171) as a device reduction, and moving bodies
(solid_dynamics.c:37-48: kinematics, collisions, motion, the
reconstruction of uncovered nodes) against the reference's final body
states."""
import hashlib
import os
import re

import numpy as np
import pytest

from conftest import case_dir, golden, rel_l2

pytestmark = pytest.mark.gpu

NUM = re.compile(r"[-+]?(?:\d+\.?\d*(?:[eE][-+]?\d+)?|inf|nan)")


def _text_files(fx):
    names = [n.decode() for n in fx["text_names"]]
    off = fx["text_offsets"]
    blob = fx["text_blob"].tobytes()
    return {n: blob[off[q]:off[q + 1]] for q, n in enumerate(names)}


def _numbers(text):
    return np.array([float(v) for v in NUM.findall(text.decode())])


def _solve(name, tmp_path, exact):
    from artracfd_amd import Case, Solver
    out = tmp_path / ("exact" if exact else "fast")
    out.mkdir()
    here = os.getcwd()
    case = Case(case_dir(name))
    solver = Solver(case, exact=exact)
    os.chdir(out)
    try:
        solver.solve()
    finally:
        os.chdir(here)
    U = case.interior(solver.field()).copy()
    F = case.forces()
    B = case.bodies()
    now, steps = case.clock()
    solver.close()
    case.close()
    return out, U, F, B, now, steps


# inviscid cases: exact arithmetic reproduces the reference bit for bit
BITWISE = ("cyl64_probes", "cyl64_ensight", "cyl64_moving", "ico_moving", "spheres_collide")


@pytest.mark.parametrize("name", ["cyl64_probes", "sphere20_probes", "cyl64_ensight", "sphere20_ensight",
                                  "cyl64_moving", "ico_moving", "spheres_collide", "sphere20_moving"])
@pytest.mark.parametrize("exact", [True, False])
def test_solve_writes_the_reference_outputs(name, exact, tmp_path):
    """Exact arithmetic on the inviscid cases: every text file
    byte-identical, every field and geometry file the same SHA-256 as the
    reference's (ParaView and EnSight Gold writers) and, for the moving
    bodies (a translating, spinning STL cylinder; a tumbling 1280-facet STL
    sphere; two spheres that collide),
    the same final positions, velocities and accelerations.  The viscous
    spheres, still and moving (the device's pow differs from glibc's in the
    last bit), and the fast arithmetic: the same files with the same numbers
    to the printed precision."""
    print(f"solve {name} exact={exact}", flush=True)
    fx = golden("solve_" + name)
    out, U, F, B, now, steps = _solve(name, tmp_path, exact)
    assert steps == int(fx["steps"]) and now == float(fx["now"])
    bitwise = exact and name in BITWISE
    err = rel_l2(U, fx["U_final"])
    assert max(err.values()) <= (0.0 if bitwise else 1e-10), err
    if bitwise:
        assert np.array_equal(F, fx["force"])
        assert np.array_equal(B, fx["bodies"])
    else:
        np.testing.assert_allclose(F, fx["force"], rtol=1e-8, atol=1e-12 * np.abs(fx["force"]).max())
        np.testing.assert_allclose(B, fx["bodies"], rtol=1e-8, atol=1e-12 * np.abs(fx["bodies"]).max())
    texts = _text_files(fx)
    written = set(os.listdir(out))
    hashed = [n.decode() for n in fx["hash_names"]]
    assert set(texts) | set(hashed) == written, (sorted(written), sorted(set(texts) | set(hashed)))
    for fname, ref in texts.items():
        got = open(out / fname, "rb").read()
        if bitwise or got == ref:
            assert got == ref, fname
            continue
        a, b = _numbers(got), _numbers(ref)
        assert a.shape == b.shape, fname
        # "%.6g" numbers: one unit in the sixth digit, or 1e-10 absolute for
        # the near-zero values (velocity components that cancel) where the
        # field's own 1e-10 bound decides the printed digits
        np.testing.assert_allclose(a, b, rtol=2e-6, atol=1e-10, err_msg=fname)
    for fname, sha, size, lines in zip(hashed, fx["hash_sha256"], fx["hash_sizes"], fx["hash_lines"]):
        data = open(out / fname, "rb").read()
        if bitwise:
            assert hashlib.sha256(data).hexdigest() == sha.decode(), fname
        elif fname.endswith((".vts", ".vtp")):
            # "%.6g" text: the same lines, lengths that vary with the last digits
            assert data.count(b"\n") == int(lines), fname
            assert abs(len(data) - int(size)) <= 0.05 * int(size), fname
        else:
            # EnSight binary: fixed-size records
            assert len(data) == int(size), fname


def test_surface_force_call_matches_the_oracle(oracle_mod):
    """artra_hip_surface_force on a state the oracle also holds (the
    cylinder after 12 steps, exact arithmetic): the same forces bit for bit."""
    from artracfd_amd import Case, Solver
    print("surface force cyl64_probes 12 steps", flush=True)
    ref = Case(case_dir("cyl64_probes"))
    oracle_mod.treat_boundary(ref, 0)
    oracle_mod.march(ref, 12)
    Fo = oracle_mod.surface_force(ref)
    ref.close()
    case = Case(case_dir("cyl64_probes"))
    solver = Solver(case, exact=True)
    solver.march(12)
    solver.surface_force()
    F = case.forces()
    solver.close()
    case.close()
    assert np.array_equal(F, Fo)
    assert np.abs(F[0, :2]).max() > 0.0


@pytest.mark.parametrize("name,first,second", [("cyl64", 15, 25), ("shockvortex64", 6, 9), ("tgv16", 7, 8),
                                               ("spheres_collide", 3, 5), ("ico_moving", 7, 9)])
def test_restart_continues_the_run_bit_for_bit(name, first, second, tmp_path):
    """March `first` steps, write the FP64 restart, continue `second` steps;
    a fresh solver reading the restart and marching `second` steps ends
    with the same field, clock, dt sequence and bodies bit for bit
    (immersed body, fused sweeps, viscous flux, colliding moving bodies)."""
    from artracfd_amd import Case, Solver
    print(f"restart {name} {first}+{second}", flush=True)
    path = str(tmp_path / "restart.bin")
    case = Case(case_dir(name))
    solver = Solver(case)
    solver.march(first)
    solver.write_restart(path)
    _, dts_a = solver.march(second)
    Ua = case.interior(solver.field()).copy()
    clock_a = case.clock()
    bodies_a = case.bodies()
    solver.close()
    case.close()
    case = Case(case_dir(name))
    solver = Solver(case)
    solver.read_restart(path)
    assert case.clock()[1] == first
    _, dts_b = solver.march(second)
    Ub = case.interior(solver.field()).copy()
    clock_b = case.clock()
    bodies_b = case.bodies()
    solver.close()
    case.close()
    assert clock_b == clock_a
    assert np.array_equal(dts_b, dts_a)
    assert np.array_equal(Ub, Ua)
    assert np.array_equal(bodies_b, bodies_a)


def test_restart_refuses_another_grid(tmp_path):
    from artracfd_amd import ArtraError, Case, Solver
    path = str(tmp_path / "restart.bin")
    case = Case(case_dir("cyl64"))
    solver = Solver(case)
    solver.write_restart(path)
    solver.close()
    case.close()
    case = Case(case_dir("sod"))
    solver = Solver(case)
    with pytest.raises(ArtraError):
        solver.read_restart(path)
    solver.close()
    case.close()
