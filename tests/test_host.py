"""Host side of libartra_hip on the CPU: the exported C-ABI, case loading,
the expression calculator and the static tables, checked against the
reference's golden fixtures and against hand-computed answers."""
import ctypes
import os
import re
import shutil

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, case_dir, golden

ALL_CASES = ["sod", "tgv16", "cyl64", "sphere20", "v_base", "v_weno3", "v_rk2", "v_roe", "v_sw", "v_unsplit",
             "v_mix", "v_visc", "v_collapse", "v_walls", "v_gravity", "v_gravity2"]


def declared_functions(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from artracfd_amd import library
    lib = ctypes.CDLL(library()._name)
    header = os.path.join(ROOT, "include", "artra_hip.h")
    names = declared_functions(header)
    assert "EvolveFluidDynamics" in names and "TreatBoundary" in names and "artra_hip_attach" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("name", ALL_CASES)
def test_prepared_case_matches_reference_flags_and_field(name):
    """Node classification (did, fid, lid, gst) over the whole padded grid,
    and the initial field on every interior node the boundary treatment does
    not overwrite (fluid nodes and body nodes outside the ghost layers)."""
    from artracfd_amd import Case
    g = golden(name)
    case = Case(case_dir(name))
    assert case.n == tuple(int(v) for v in g["n"])
    assert np.array_equal(case.flags(), g["flags"].astype(np.int32))
    fl = case.interior(case.flags())
    keep = (fl[0] == 0) | ((fl[0] > 0) & (fl[3] == 0))
    assert np.array_equal(case.interior(case.field())[:, keep], g["U_init"][:, keep])
    case.close()


@pytest.mark.parametrize("expr, value", [
    ("1.5*sin(-pi/6)-[cos(pi/3)]^2+ln{exp[5*lg(abs(-100))]}", 9.0),
    ("2^3^2", 64.0),
    ("-2^2", -4.0),
    ("2*3+4*5", 26.0),
    ("8/4/2", 1.0),
    ("sin(0.5)^2", np.sin(0.5) ** 2),
    ("(1+2)*{3-[4/(5+3)]}", 7.5),
    ("tan(1)", np.tan(1.0)),
    ("abs(-3)+exp(0)", 4.0),
])
def test_calculator_grammar(expr, value):
    from artracfd_amd.lib import calc
    assert calc(expr) == pytest.approx(value, rel=1e-15, abs=1e-15)


def test_calculator_variables_and_errors():
    from artracfd_amd.lib import calc
    assert calc("x*y+z", 2.0, 3.0, 4.0) == 10.0
    assert calc("t+1") == 1.0
    for bad in ["1/0", "ln(0)", "lg(-1)", "2*(3", "2+", "abc", ".5", "5.", "1e-3", "2**3", "(1))"]:
        assert calc(bad) is None, bad


def _write_case(tmp_path, edit):
    src = case_dir("sod")
    dst = tmp_path / "case"
    shutil.copytree(src, dst)
    text = open(dst / "artracfd.case").read()
    open(dst / "artracfd.case", "w").write(edit(text))
    return dst


@pytest.mark.parametrize("edit, message", [
    (lambda t: t.replace("space begin", "spaces begin"), "missing or repeated sections"),
    (lambda t: t.replace("outflow", "outflux", 1), "unidentified boundary type"),
    (lambda t: t.replace("256, 4, 4", "256, 0, 4"), "mesh number should be positive"),
    (lambda t: t.replace("time begin\n0\n0.2", "time begin\n0\n-0.2"), "time section"),
    (lambda t: t.replace("reference begin\n1.0", "reference begin\n0.0"), "reference values"),
    (lambda t: t.replace("numerical begin\n1\n1", "numerical begin\n1\n7"), "numerical section"),
])
def test_case_errors_are_reported(tmp_path, edit, message):
    from artracfd_amd import ArtraError, Case
    d = _write_case(tmp_path, edit)
    with pytest.raises(ArtraError) as err:
        Case(d)
    assert message in str(err.value)


def test_missing_case_directory_is_an_error(tmp_path):
    from artracfd_amd import ArtraError, Case
    with pytest.raises(ArtraError) as err:
        Case(tmp_path / "nothing")
    assert "failed to open file" in str(err.value)


def test_expression_too_long_is_an_error(tmp_path):
    from artracfd_amd import ArtraError, Case
    d = _write_case(tmp_path, lambda t: t.replace("initialization begin\n0.125", "initialization begin\n" + "1+" * 60 + "1"))
    with pytest.raises(ArtraError):
        Case(d)


def test_geometry_of_the_cylinder():
    """The 2048-facet cylinder: closed surface (every edge shared by two
    facets), volume and bounding box of the analytic cylinder."""
    from artracfd_amd import Case
    case = Case(case_dir("cyl64"))
    fl = case.flags()
    inside = (fl[0] == 1)
    # nodes marked inside lie within the radius (plus the facet chord sag)
    n = case.n
    d = 1.0 / 64
    k, j, i = np.nonzero(inside)
    x = 0.5 + (i - 2) * d
    y = 0.5 + (j - 2) * d
    r = np.hypot(x - 1.0, y - 1.0)
    assert r.max() <= 0.1 + 1e-9 and inside.sum() > 0
    # every node within 0.099 of the axis (and in the interior range) is inside
    kk, jj, ii = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    rr = np.hypot(0.5 + (ii - 2) * d - 1.0, 0.5 + (jj - 2) * d - 1.0)
    deep = case.interior(rr < 0.099)
    assert np.all(case.interior(inside)[deep])
    # ghost layers 1..3 exist and sit inside the body
    gst = case.interior(fl[3])
    assert set(np.unique(gst[case.interior(inside)])) >= {1, 2, 3}
    case.close()


def test_drop_in_structures_have_the_reference_layout(tmp_path):
    """sizeof and offsetof of every member of Node, Partition, Facet,
    Collision, Polyhedron, Geometry, Material, Space, Time and Model in
    include/artra_types.h equal those of the reference's commons.h (dumped by
    the reference harness into tests/golden/abi_layout.txt)."""
    import subprocess
    exe = str(tmp_path / "abi_layout")
    subprocess.run(["gcc", "-std=c11", "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "oracle"),
                    os.path.join(ROOT, "tests", "abi_layout.c"), "-o", exe], check=True)
    ours = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.splitlines()
    ref = open(os.path.join(GOLDEN, "abi_layout.txt")).read().splitlines()
    assert len(ref) == 118 and ours == ref


REF_OBJ = os.path.join(ROOT, "oracle", "_ref", "obj")
REPLACED = ["fluid_dynamics", "boundary_treatment", "immersed_boundary", "convective_flux", "diffusive_flux",
            "weno3", "weno5", "source_term"]


def _symbols(path, *flags):
    import subprocess
    out = subprocess.run(["nm", *flags, path], check=True, capture_output=True, text=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


@pytest.mark.skipif(not os.path.isdir(REF_OBJ), reason="the reference objects are built only where its sources are")
def test_integration_link_recipe():
    """The link recipe of INTEGRATION.md: the reference's objects minus
    fluid_dynamics, boundary_treatment, immersed_boundary and the files only
    they call, plus libartra_hip.so, link into the reference's driver with no
    undefined symbol (oracle/Makefile, target dropin), the library defines
    every symbol the kept objects take from the dropped ones, and no symbol
    is defined twice."""
    import subprocess
    from artracfd_amd import library
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "dropin"], check=True)
    objs = sorted(f for f in os.listdir(REF_OBJ) if f.endswith(".o"))
    kept = [f for f in objs if f[:-2] not in REPLACED]
    assert len(kept) == len(objs) - len(REPLACED)
    defined_kept, undefined_kept, defined_dropped = set(), set(), set()
    for f in kept:
        defined_kept |= _symbols(os.path.join(REF_OBJ, f), "-g", "--defined-only")
        undefined_kept |= _symbols(os.path.join(REF_OBJ, f), "-u")
    for name in REPLACED:
        defined_dropped |= _symbols(os.path.join(REF_OBJ, name + ".o"), "-g", "--defined-only")
    exported = _symbols(library()._name, "-D", "--defined-only")
    needed = undefined_kept & defined_dropped
    assert needed == {"EvolveFluidDynamics", "TreatBoundary", "TreatImmersedBoundary", "ComputeGeometricField"}
    assert needed <= exported
    assert not (defined_kept & exported), defined_kept & exported


@pytest.mark.parametrize("name, moves, motion", [
    ("cyl64_moving", 12, ("0.004", "0.0025", "0", "0.02")),     # sub-cell steps with spin
    ("cyl64_moving", 5, ("0.031", "-0.02", "0", "0.3")),         # two cells a move
    ("ico_moving", 6, ("0.011", "-0.007", "0.005", "0.25")),     # a 1280-facet sphere, 3D facets
    ("spheres_collide", 10, ("0.012", "-0.006", "0.003", "0")),
    ("spheres_collide", 6, ("-0.05", "0.03", "0.02", "0")),
])
def test_moving_body_classification_shortcuts(name, moves, motion, tmp_path):
    """After each move of the bodies, the classification restricted to the
    scope's region (where the moving bodies were or are, grown by the ghost
    layers, and the last reconstructed nodes) leaves every node's did, fid,
    lid and gst and the reconstruction lists exactly as the classification
    of the whole interior (immersed_boundary.c:65-264); the bounded
    closest-facet search picks the same facet as the full scan on every
    interior node and on vertices, edge midpoints and facet centroids; the
    ghost table built in parallel equals the serial one (tests/scope_check.c)."""
    import subprocess
    from artracfd_amd import library
    lib = library()._name
    exe = str(tmp_path / "scope_check")
    subprocess.run(["gcc", "-O2", "-std=c11", "-fopenmp", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "artracfd_amd", "csrc", "host"),
                    os.path.join(ROOT, "tests", "scope_check.c"), "-o", exe, lib,
                    "-Wl,-rpath," + os.path.dirname(lib), "-lm"], check=True)
    out = subprocess.run([exe, case_dir(name), str(moves), *motion], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok ") and int(out.stdout.split()[1]) > 0, out.stdout
