#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden from the compiled reference.

Runs oracle/_ref/artra_ref (the reference solver's own sources compiled by
oracle/Makefile, only possible where those sources exist) on each small case
and on seeded per-function inputs, and stores inputs and outputs as
compressed numpy archives:

  <case>.npz   dt[steps], steps, now, flags[4][nz][ny][nx] (int16), and the
               interior (ns[PIN]) slice of level TO after initialisation
               (U_init) and after the run (U_final), FP64
  vectors.npz  WENO5/WENO3, face fluxes on a random grid, primitive maps,
               point-triangle distances
  abi_layout.txt  sizeof/offsetof of the drop-in structures (commons.h)
  solve_<case>.npz  the reference's whole Solve on a case with probes: the
               text of every probe and log file it writes, the SHA-256 and
               size of its field and geometry files, the final surface
               force of each body (Fp, Fv, Tt), its final motion state (O,
               V, W, at, ar) and the final interior field

    python tests/golden/make_golden.py [case ...]
"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

# case -> steps (0: the case's own end time)
CASES = {
    "sod": 0, "tgv16": 20, "cyl64": 60, "sphere20": 20,
    "v_base": 10, "v_weno3": 10, "v_rk2": 10, "v_roe": 10, "v_sw": 10, "v_unsplit": 10,
    "v_mix": 10, "v_visc": 10, "v_collapse": 10, "v_walls": 10, "v_gravity": 10, "v_gravity2": 10,
}


def interior(arr, ng_pin):
    (x0, x1), (y0, y1), (z0, z1) = ng_pin
    return arr[..., z0:z1, y0:y1, x0:x1]


def pin_of(case_dir):
    sys.path.insert(0, ROOT)
    from artracfd_amd import Case
    c = Case(case_dir)
    pin = c.pin
    c.close()
    return pin


def make_case(name, steps, work):
    case_dir = os.path.join(ROOT, "cases", name)
    info, paths = oracle.run_reference(case_dir, os.path.join(work, name), steps)
    hi, fi, Ui = oracle.read_dump(paths["init"])
    hf, ff, Uf = oracle.read_dump(paths["final"])
    dts = np.fromfile(paths["dt"])
    pin = pin_of(case_dir)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), dt=dts, steps=info["steps"], now=info["now"],
                        flags=fi.astype(np.int16), pin=np.array(pin), n=np.array(hi["n"]),
                        U_init=np.ascontiguousarray(interior(Ui, pin)),
                        U_final=np.ascontiguousarray(interior(Uf, pin)))
    print(name, info)


def make_vectors(work, seed=20261202, count=400):
    prefix = os.path.join(work, "vec")
    import subprocess
    subprocess.run([oracle.REF, "vectors", prefix, str(seed), str(count)], check=True, capture_output=True)
    rd = lambda n, dt=np.float64: np.fromfile(f"{prefix}_{n}.bin", dtype=dt)
    nf = int(rd("fhat_nfaces", np.int32)[0])
    np.savez_compressed(
        os.path.join(HERE, "vectors.npz"),
        weno5_in=rd("weno5_in").reshape(count, 5, 5), weno5_out=rd("weno5_out").reshape(count, 5),
        weno3_in=rd("weno3_in").reshape(count, 3, 5), weno3_out=rd("weno3_out").reshape(count, 5),
        grid_U=rd("grid_U").reshape(10, 10, 10, 5), faces=rd("fhat_faces", np.int32).reshape(nf, 3),
        fhat_out=rd("fhat_out").reshape(8, 3, nf, 5), fvhat_out=rd("fvhat_out").reshape(3, nf, 5),
        fv_params=np.array([1.4, 2.5, 0.37, 1.7, 1 / 0.1, 1 / 0.07, 1 / 0.13]),
        prim_in=rd("prim_in").reshape(count, 5), prim_out=rd("prim_out").reshape(count, 6),
        cons_out=rd("cons_out").reshape(count, 5), prim_gasR=np.array([1.0 / (1.4 * 0.1 * 0.1)]),
        ptd_in=rd("ptd_in").reshape(count, 12), ptd_out=rd("ptd_out").reshape(count, 4))
    print("vectors", count, nf)


SOLVE_CASES = ("cyl64_probes", "sphere20_probes", "cyl64_ensight", "sphere20_ensight", "cyl64_moving", "ico_moving",
               "spheres_collide", "sphere20_moving")
SOLVE_TEXT = (".csv", ".log", ".pvd", ".case", ".state")   # stored as text; other outputs (.vts, .vtp, EnSight binaries) by hash


def make_solve(name, work):
    """Every output file of the reference's Solve on the case (run in a copy
    of the case directory), with the inputs and the harness's dumps left
    out."""
    import hashlib
    import json
    import shutil
    import subprocess
    case_dir = os.path.join(ROOT, "cases", name)
    d = os.path.join(work, "solve_" + name)
    shutil.copytree(case_dir, d)
    inputs = set(os.listdir(d))
    out = subprocess.run([oracle.REF, "solve", "out"], cwd=d, check=True, capture_output=True, text=True)
    info = json.loads(out.stdout.strip().splitlines()[-1])
    names = sorted(f for f in os.listdir(d) if f not in inputs and not f.startswith("out_")
                   and f != "artracfd.verify" and f != "geo_mass_property.csv")
    texts, hashed = [], []
    for f in names:
        data = open(os.path.join(d, f), "rb").read()
        (texts if f.endswith(SOLVE_TEXT) else hashed).append((f, data))
    blob = b"".join(data for _, data in texts)
    offsets = np.cumsum([0] + [len(data) for _, data in texts])
    hf, ff, Uf = oracle.read_dump(os.path.join(d, "out_final.bin"))
    pin = pin_of(case_dir)
    np.savez_compressed(os.path.join(HERE, f"solve_{name}.npz"),
                        text_names=np.array([f for f, _ in texts], dtype="S64"),
                        text_offsets=offsets.astype(np.int64),
                        text_blob=np.frombuffer(blob, dtype=np.uint8),
                        hash_names=np.array([f for f, _ in hashed], dtype="S64"),
                        hash_sha256=np.array([hashlib.sha256(data).hexdigest() for _, data in hashed], dtype="S64"),
                        hash_sizes=np.array([len(data) for _, data in hashed], dtype=np.int64),
                        hash_lines=np.array([data.count(b"\n") for _, data in hashed], dtype=np.int64),
                        force=np.fromfile(os.path.join(d, "out_force.bin")).reshape(-1, 9),
                        bodies=np.fromfile(os.path.join(d, "out_bodies.bin")).reshape(-1, 15),
                        steps=info["steps"], now=info["now"],
                        U_final=np.ascontiguousarray(interior(Uf, pin)))
    print(name, info, len(texts), "text files", len(hashed), "hashed files")


def make_layout():
    """sizeof/offsetof of the drop-in structures as the reference's own
    commons.h lays them out (the harness's layout dump)."""
    import subprocess
    out = subprocess.run([oracle.REF, "layout"], check=True, capture_output=True, text=True).stdout
    with open(os.path.join(HERE, "abi_layout.txt"), "w") as fp:
        fp.write(out)
    print("layout", len(out.splitlines()), "entries")


def main(argv):
    if not os.path.exists(oracle.REF):
        oracle.build()
    names = argv or list(CASES)
    with tempfile.TemporaryDirectory() as work:
        for name in names:
            if name == "vectors":
                make_vectors(work)
            elif name == "layout":
                make_layout()
            elif name in SOLVE_CASES:
                make_solve(name, work)
            else:
                make_case(name, CASES[name], work)
        if not argv:
            make_vectors(work)
            make_layout()
            for name in SOLVE_CASES:
                make_solve(name, work)


if __name__ == "__main__":
    main(sys.argv[1:])
