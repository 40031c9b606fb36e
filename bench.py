#!/usr/bin/env python3
"""Benchmark of the fluid-update hot path on one MI355X.

Workload: the 256^3 shock-vortex interaction (cases/shockvortex256: Mach 1.5
shock at x = 0.5, Gaussian isentropic vortex of strength 0.3 and radius 0.075
at (0.75, 1, 1), domain [0,2]^3, inviscid, WENO5 + LLF + RK3 with Strang
splitting, CFL 0.6), plain FP64 path.  A step is one full solver time step:
the device CFL reduction and dt readback, 18 directional sweeps and 18
boundary fills.  The metric counts cell-stage updates: interior fluid nodes x
18 x steps / wall seconds.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--case NAME] [--no-cpu]

Prints one JSON line.  The roofline object is for the dominant kernel family
(the directional sweeps), from HIP events on the solver's stream over the
timed region; algorithmic bytes are 80 B per cell-stage for the first RK3
stage (read U_tn, write U_tm) and 120 B for the other two (also U_to).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# Measured once with rocprofv3 on this workload (profiles/sweep_shockvortex256.md):
# HBM bytes per cell-stage of the sweep kernels (FETCH_SIZE x 2.0 + WRITE_SIZE,
# scale factors from a 1 GiB copy kernel) and FP64 flops per face (FMA = 2).
TRAFFIC_B_PER_CELL_STAGE = 129.6
FP64_FLOP_PER_CELL_STAGE = 1124.0
METRIC = "Mcell-stage updates/s (256³ WENO5-RK3) + learned-hybrid vs plain at ≤1e-3 L1"


def fluid_nodes(case):
    flags = case.flags()
    return int((case.interior(flags[0]) == 0).sum())


def stage_bytes(tscheme):
    # RK3: stage 1 reads U_tn and writes U_tm; stages 2 and 3 also read U_to
    return (80.0 + 120.0 + 120.0) / 3.0 if tscheme == 1 else (80.0 + 120.0) / 2.0


def cpu_baseline(seconds_budget=12.0):
    """The CPU oracle (single-thread restatement of the reference's hot path,
    FP64, -O2) on the 64^3 version of the same shock-vortex case."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from artracfd_amd import Case
    case = Case(os.path.join(ROOT, "cases", "shockvortex64"))
    oracle.treat_boundary(case, 0)
    cells = fluid_nodes(case)
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds_budget:
        n, _ = oracle.march(case, 1)
        steps += n
        if n == 0:
            break
    el = time.perf_counter() - t0
    case.close()
    return {"value": cells * 18 * steps / el / 1e6, "unit": "Mcell-stage updates/s", "cores": 1, "kind": "port",
            "sample": f"shockvortex64 (63^3 interior nodes), {steps} full time steps, {el:.1f} s, "
                      f"oracle/artra_oracle.c single thread"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--case", default="shockvortex256")
    ap.add_argument("--segment", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the multi-GPU slab decomposition is a "
                                   "design (see DESIGN.md), so --gpus must be 1"}))
        sys.exit(2)
    from artracfd_amd import Case, Solver
    t_init = time.perf_counter()
    case = Case(os.path.join(ROOT, "cases", args.case))
    cells = fluid_nodes(case)
    solver = Solver(case, segment=args.segment)
    init_s = time.perf_counter() - t_init
    from artracfd_amd.lib import copy_probe
    copy_gbs = copy_probe(1 << 30, 20)
    solver.march(args.warmup)
    solver.synchronize()
    solver.set_timing(True)
    t0 = time.perf_counter()
    done, dts = solver.march(args.steps)
    solver.synchronize()
    elapsed = time.perf_counter() - t0
    tm = solver.timing()
    value = cells * 18 * done / elapsed / 1e6
    sweep_ms, sweep_n = tm["sweep"]
    bytes_per_launch = cells * stage_bytes(1)
    avg_launch_s = sweep_ms / 1e3 / max(sweep_n, 1)
    achieved = bytes_per_launch / avg_launch_s / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcell-stage updates/s",
        "n_gpus": 1,
        "steps": done,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / max(done, 1), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.case}: shock-vortex interaction, {cells} interior fluid nodes, WENO5-LLF-RK3 "
                               f"split, inviscid, plain FP64 path", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(TRAFFIC_B_PER_CELL_STAGE * cells) if args.case == "shockvortex256" else None,
                     "kernel": "sweep_x/sweep_march (directional WENO5 sweeps)",
                     "avg_launch_ms": round(avg_launch_s * 1e3, 4), "launches": sweep_n,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "measured_copy_gbs": round(copy_gbs, 1),
                     "fp64": {"achieved_tflops": round(FP64_FLOP_PER_CELL_STAGE * cells / avg_launch_s / 1e12, 2),
                              "peak_tflops": FP64_PEAK_TFLOPS,
                              "frac": round(FP64_FLOP_PER_CELL_STAGE * cells / avg_launch_s / 1e12 / FP64_PEAK_TFLOPS, 4),
                              "flop_per_cell_stage": FP64_FLOP_PER_CELL_STAGE}},
        "kernel_ms": {k: round(v[0], 3) for k, v in tm.items()},
        "kernel_launches": {k: v[1] for k, v in tm.items()},
        "init_seconds": round(init_s, 2),
    }
    solver.close()
    case.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
