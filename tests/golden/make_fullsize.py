#!/usr/bin/env python3
"""Fixtures of the full-size reference runs (TGV 128^3 after 200 steps,
shock-cylinder 512x512x4 after 1000 steps), which take hours on one CPU
core and are too large to commit whole.

Reads the FP64 dumps written by oracle/_ref/artra_ref for a case directory
(`artra_ref run 0 out` in a copy of cases/<name>) and stores

  <name>_full.npz  dt[steps], steps, now, pin; per component the sum and the
                   sum of squares of level TO over the interior fluid nodes;
                   a strided subsample of the interior (every `stride`-th
                   node along each axis); the SHA-256 of the interior array

    python tests/golden/make_fullsize.py <name> <run-directory> [stride]
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    name, run = sys.argv[1], sys.argv[2]
    stride = int(sys.argv[3]) if len(sys.argv) > 3 else 8
    from artracfd_amd import Case
    case = Case(os.path.join(ROOT, "cases", name))
    pin = case.pin
    case.close()
    hdr, flags, U = oracle.read_dump(os.path.join(run, "out_final.bin"))
    dts = np.fromfile(os.path.join(run, "out_dt.bin"))
    (x0, x1), (y0, y1), (z0, z1) = pin
    Ui = np.ascontiguousarray(U[:, z0:z1, y0:y1, x0:x1])
    fluid = flags[0, z0:z1, y0:y1, x0:x1] == 0
    sums = np.array([Ui[c][fluid].sum() for c in range(5)])
    squares = np.array([(Ui[c][fluid] ** 2).sum() for c in range(5)])
    sub = np.ascontiguousarray(Ui[:, ::stride, ::stride, ::stride])
    digest = hashlib.sha256(Ui.tobytes()).hexdigest()
    np.savez_compressed(os.path.join(HERE, f"{name}_full.npz"), dt=dts, steps=hdr["steps"], now=hdr["now"],
                        pin=np.array(pin), stride=stride, sums=sums, squares=squares, sub=sub,
                        sha256=np.array(digest))
    print(name, hdr, "subsample", sub.shape, digest)


if __name__ == "__main__":
    main()
