"""Build libartra_hip.so in-tree.

Host side: C11 with gcc (OpenMP for the initial-field loop).  Device side:
HIP C++ for gfx950 with hipcc.  Both use -ffp-contract=off so the FP64
arithmetic keeps the solver's operation order.  Packed FP32 instructions are
disabled for all device code.

    python -m artracfd_amd.build        # or artracfd_amd.build.build()
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libartra_hip.so")
ARCH = os.environ.get("ARTRA_OFFLOAD_ARCH", "gfx950")

HOST_SOURCES = ["case_io.c", "calculator.c", "setup.c", "geometry.c", "tables.c", "capi.c", "output.c", "motion.c"]
DEVICE_SOURCES = ["artra_device.hip", "indicator.hip"]

HOST_FLAGS = ["-O2", "-std=c11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
              "-Wall", "-Wextra", "-Wno-unused-parameter"]
DEVICE_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off", "-fno-fast-math",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return proc.stdout


def build(verbose=False):
    """Compile every source and link libartra_hip.so (always rebuilds)."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    os.makedirs(OBJ, exist_ok=True)
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(CSRC, "host"),
           "-I" + os.path.join(CSRC, "device")]
    objs = []
    for src in HOST_SOURCES:
        out = os.path.join(OBJ, src + ".o")
        _run(["gcc"] + HOST_FLAGS + inc + ["-c", os.path.join(CSRC, "host", src), "-o", out])
        objs.append(out)
    for src in DEVICE_SOURCES:
        out = os.path.join(OBJ, src + ".o")
        _run([hipcc] + DEVICE_FLAGS + inc + ["-c", os.path.join(CSRC, "device", src), "-o", out])
        objs.append(out)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-fopenmp", "-lm"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
