#!/usr/bin/env python3
"""Benchmark of the fluid-update hot path on one MI355X.

Workload: the 256^3 shock-vortex interaction (cases/shockvortex256: Mach 1.5
shock at x = 0.5, Gaussian isentropic vortex of strength 0.3 and radius 0.075
at (0.75, 1, 1), domain [0,2]^3, inviscid, WENO5 + LLF + RK3 with Strang
splitting, CFL 0.6), plain FP64 path.  A step is one full solver time step:
the device CFL reduction and dt readback, six directional operators of three
RK stages each, and their boundary fills.  The metric counts cell-stage
updates: interior fluid nodes x 18 x steps / wall seconds.

The learned-hybrid path (committed indicator weights) is timed the same way
on the same workload, and both paths are run to t = 0.6 to report the
relative L1 density difference of the hybrid against plain WENO5.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--case NAME] [--no-cpu] [--no-hybrid]

Prints one JSON line.  The roofline object is for the dominant kernel family
(the directional sweeps), from HIP events on the solver's stream over the
timed region.  Algorithmic bytes per cell-stage are those of a stage-by-stage
sweep (80 B for the first RK3 stage: read U_tn, write U_tm; 120 B for the
other two, which also read U_to; 106.7 B on average); a fused launch covers
three cell-stages per cell.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# Measured with rocprofv3 on this workload (profiles/rk_line_shockvortex256.md):
# HBM bytes per cell-stage of the fused sweep kernel (FETCH_SIZE x 2.0 +
# WRITE_SIZE, scale factors from a 1 GiB copy kernel) and FP64 flops per
# cell-stage of the plain WENO5 flux and update (FMA = 2, from the FP64
# instruction counters).
TRAFFIC_B_PER_CELL_STAGE = 25.9
FP64_FLOP_PER_CELL_STAGE = 914.0
# the stage-by-stage sweeps (profiles/sweep_shockvortex256.md)
UNFUSED_TRAFFIC_B_PER_CELL_STAGE = 112.3
METRIC = "Mcell-stage updates/s (256³ WENO5-RK3) + learned-hybrid vs plain at ≤1e-3 L1"


def fluid_nodes(case):
    flags = case.flags()
    return int((case.interior(flags[0]) == 0).sum())


WORKLOADS = {
    "shockvortex256": "shock-vortex interaction",
    "shockvortex128": "shock-vortex interaction",
    "shockvortex64": "shock-vortex interaction",
    "tgv128": "viscous Taylor-Green vortex",
    "cylinder": "shock diffraction over a cylinder (immersed boundary)",
    "cylinder_moving": "shock over a translating, spinning cylinder (moving immersed boundary, solid dynamics)",
    "sod": "Sod shock tube",
}


def describe(case_dir, name, cells):
    """Workload line: the case, its size and its scheme as its case file states it."""
    lines = [ln.split("#")[0].strip() for ln in open(os.path.join(case_dir, "artracfd.case"))]
    lines = [ln for ln in lines if ln]
    num = lines[lines.index("numerical begin") + 1:]
    mat = lines[lines.index("material begin") + 1:]
    rk = "RK3" if num[0] == "1" else "RK2"
    weno = "WENO5" if num[1] == "1" else "WENO3"
    split = "split" if num[2] == "0" else "unsplit"
    flux = ("Roe-" if num[3] == "1" else "") + ("SW" if num[4] == "1" else "LLF")
    visc = "viscous" if float(mat[1]) != 0.0 else "inviscid"
    what = WORKLOADS.get(name, "case")
    return f"{name}: {what}, {cells} interior fluid nodes, {weno}-{flux}-{rk} {split}, {visc}, plain FP64 path"


def stage_bytes(tscheme):
    # RK3: stage 1 reads U_tn and writes U_tm; stages 2 and 3 also read U_to
    return (80.0 + 120.0 + 120.0) / 3.0 if tscheme == 1 else (80.0 + 120.0) / 2.0


def cpu_baseline(steps=10):
    """CPU rate of the same shock-vortex case at 64^3 (63^3 interior nodes),
    `steps` full time steps on one thread (the reference has no threading):
    the reference's own solver where its compiled harness is present
    (oracle/_ref/artra_ref, built from the reference sources in the build
    container), and the CPU oracle (oracle/artra_oracle.c, the single-thread
    restatement of the hot path, FP64, -O2) on the same steps in every case,
    so that the restatement's speed relative to the reference is measured on
    the same host whenever both are present."""
    import shutil
    import subprocess
    import tempfile
    src = os.path.join(ROOT, "cases", "shockvortex64")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from artracfd_amd import Case
    case = Case(src)
    oracle.treat_boundary(case, 0)
    cells = fluid_nodes(case)
    t0 = time.perf_counter()
    n, _ = oracle.march(case, steps)
    port_s = time.perf_counter() - t0
    case.close()
    port = cells * 18 * n / port_s / 1e6
    try:
        cores_avail = len(os.sched_getaffinity(0))
    except AttributeError:
        cores_avail = os.cpu_count()
    out = {"value": port, "unit": "Mcell-stage updates/s", "cores": 1, "kind": "port",
           "sample": f"shockvortex64 (63^3 interior nodes), {n} full time steps, {port_s:.1f} s, "
                     f"oracle/artra_oracle.c single thread",
           "host_cores": {"total": os.cpu_count(), "available_to_this_process": cores_avail},
           "port": {"value": round(port, 4), "seconds": round(port_s, 2), "steps": n}}
    ref = os.path.join(ROOT, "oracle", "_ref", "artra_ref")
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as tmp:
            work = os.path.join(tmp, "case")
            shutil.copytree(src, work)
            run = subprocess.run([ref, "run", str(steps), "out"], cwd=work, check=True, capture_output=True,
                                 text=True, timeout=600)
            info = json.loads(run.stdout.strip().splitlines()[-1])
        refv = info["fluid_nodes"] * 18 * info["steps"] / info["seconds"] / 1e6
        out.update({"value": refv, "kind": "reference",
                    "sample": f"shockvortex64 (63^3 interior nodes), {info['steps']} full time steps, "
                              f"{info['seconds']:.1f} s, the reference's solver compiled from its own sources "
                              f"(oracle/_ref), single thread",
                    "reference": {"value": round(refv, 4), "seconds": round(info["seconds"], 2),
                                  "steps": info["steps"]},
                    "port_over_reference": round(port / refv, 3)})
    return out


def reference_rate(name, steps):
    """The reference's own solver (oracle/_ref/artra_ref, compiled from its
    sources in the build container) on cases/<name> for `steps` full time
    steps, single thread: (Mcell-stage updates/s, seconds, steps), or None
    where the harness is absent."""
    import shutil
    import subprocess
    import tempfile
    ref = os.path.join(ROOT, "oracle", "_ref", "artra_ref")
    if not os.path.exists(ref):
        return None
    with tempfile.TemporaryDirectory() as tmp:
        work = os.path.join(tmp, "case")
        shutil.copytree(os.path.join(ROOT, "cases", name), work)
        run = subprocess.run([ref, "run", str(steps), "out"], cwd=work, check=True, capture_output=True, text=True,
                             timeout=600)
        info = json.loads(run.stdout.strip().splitlines()[-1])
    return info["fluid_nodes"] * 18 * info["steps"] / info["seconds"] / 1e6, info["seconds"], info["steps"]


# The viscous and immersed-boundary workloads timed beside the headline:
# (case, timed steps, CPU sample case, CPU sample steps, dominant kernel family,
# measured HBM bytes per cell-stage of its sweeps, profiles/sweep_tgv128_cylinder.md)
SIDE_WORKLOADS = (
    ("tgv128", 20, "tgv64", 4, "sweep_x/sweep_march with the viscous flux (stage-by-stage sweeps)", 176.6),
    ("cylinder", 40, "cyl64", 100, "sweep_x/sweep_march with the immersed-boundary fill (stage-by-stage sweeps)",
     126.0),
    ("cylinder_moving", 40, "cyl64_moving", 10,
     "sweep_x/sweep_march with the immersed-boundary fill; solid half-steps (surface force, kinematics, motion, "
     "reclassification, ghost tables) between the fluid steps", 126.0),
)


def side_workloads(warmup, with_cpu):
    """Rate, kernel times and the sweep roofline of the viscous Taylor-Green
    vortex (128^3), the shock over the immersed cylinder (512x512x4) and the
    same cylinder moving (translating and spinning under the shock's force),
    with the reference's CPU rate on a smaller sample of each.  host_ms_per_step
    is the wall time not covered by kernels (solid half-steps, launches)."""
    from artracfd_amd import Case, Solver
    out = {}
    for name, steps, cpu_case, cpu_steps, kernel, traffic_b in SIDE_WORKLOADS:
        case_dir = os.path.join(ROOT, "cases", name)
        case = Case(case_dir)
        cells = fluid_nodes(case)
        solver = Solver(case)
        # the rate without kernel timing (each step replayed from its captured
        # graph), then the kernel times over the same steps of a second run
        solver.march(warmup)
        solver.synchronize()
        t0 = time.perf_counter()
        done, _ = solver.march(steps)
        solver.synchronize()
        elapsed = time.perf_counter() - t0
        solver.close()
        case.close()
        case = Case(case_dir)
        solver = Solver(case)
        tdone, telapsed, tm, _ = timed(solver, warmup, steps)
        solver.close()
        case.close()
        sweep_ms, sweep_n = tm["sweep"]
        stage_b = stage_bytes(1)
        # every sweep launch of these workloads is one RK stage of one axis
        achieved = cells * stage_b * (18.0 * tdone) / (sweep_ms / 1e3) / 1e9
        busy = {k: v[0] for k, v in tm.items()}
        entry = {
            "workload": describe(case_dir, name, cells), "value": round(cells * 18 * done / elapsed / 1e6, 3),
            "unit": "Mcell-stage updates/s", "steps": done, "ms_per_step": round(elapsed * 1e3 / max(done, 1), 4),
            "timed_pass": {"steps": tdone, "ms_per_step": round(telapsed * 1e3 / max(tdone, 1), 4),
                           "note": "kernel times from HIP events over the same steps of a second run, each kernel "
                                   "launched on its own (the rate above replays each step from its captured graph)"},
            "kernel_ms": {k: round(v, 3) for k, v in busy.items()},
            "host_ms_per_step": round(max(telapsed * 1e3 - sum(busy.values()), 0.0) / max(tdone, 1), 4),
            "kernel_launches": {k: v[1] for k, v in tm.items()},
            "dominant": max(busy, key=busy.get),
            "roofline": {"kernel": kernel, "bound": "hbm", "algorithmic_bytes_per_cell_stage": round(stage_b, 1),
                         "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic_bytes_per_cell_stage": traffic_b,
                         "avg_launch_ms": round(sweep_ms / max(sweep_n, 1), 4), "launches": sweep_n},
        }
        if with_cpu:
            r = reference_rate(cpu_case, cpu_steps)
            if r is not None:
                entry["cpu_baseline"] = {"value": round(r[0], 4), "unit": "Mcell-stage updates/s", "cores": 1,
                                         "kind": "reference",
                                         "sample": f"{cpu_case}, {r[2]} full time steps, {r[1]:.1f} s, the "
                                                   f"reference's solver compiled from its own sources, single thread"}
        out[name] = entry
    return out


def timed(solver, warmup, steps):
    solver.march(warmup)
    solver.synchronize()
    solver.set_timing(True)
    solver.hybrid_counts(reset=True)
    t0 = time.perf_counter()
    done, _ = solver.march(steps)
    solver.synchronize()
    elapsed = time.perf_counter() - t0
    tm = solver.timing()
    solver.set_timing(False)
    return done, elapsed, tm, solver.hybrid_counts()


def run_to_end(case_dir, params=None, threshold=0.0):
    """Density at the case's end time (t = 0.6) and the wall time of the march."""
    from artracfd_amd import Case, Solver
    case = Case(case_dir)
    solver = Solver(case)
    if params is not None:
        solver.set_indicator(params, threshold)
    solver.synchronize()
    t0 = time.perf_counter()
    steps, _ = solver.march()
    solver.synchronize()
    el = time.perf_counter() - t0
    rho = case.interior(solver.field())[0].copy()
    solver.close()
    case.close()
    return rho, steps, el


def batched_both(with_indicator=True):
    """The batched workload with plain WENO5, and (unless told not to) again
    with the learned indicator, reported as a sub-object."""
    plain = batched(use_indicator=False)
    if with_indicator:
        hyb = batched(use_indicator=True)
        plain["hybrid"] = {k: hyb[k] for k in ("value", "serial_value", "wall_ms", "linear_face_fraction",
                                                "weno5_face_fraction", "flagged_face_fraction", "tuned",
                                                "serial_ms_per_case")}
        plain["hybrid"]["vs_plain"] = round(hyb["value"] / plain["value"], 3)
    return plain


def batched(seed=2026, count=64, steps=100, use_indicator=True):
    """The batched workload: `count` oblique-shock cases of uneven cubic
    sizes (artracfd_amd.batch), tuned, then `steps` lockstep steps replayed
    from a HIP graph; and the same cases one at a time for per-case wall
    times."""
    import tempfile

    from artracfd_amd import Case, Solver, indicator
    from artracfd_amd.batch import Batch, oblique_shock_cases
    params = thr = None
    if use_indicator and os.path.exists(indicator.WEIGHTS):
        params, meta = indicator.load()
        thr = float(meta.get("logit_threshold", indicator.LOGIT_THRESHOLD))
    with tempfile.TemporaryDirectory() as root:
        t0 = time.perf_counter()
        dirs, drawn = oblique_shock_cases(root, seed, count, steps)
        cases = [Case(d) for d in dirs]
        cells = [fluid_nodes(c) for c in cases]
        solvers = [Solver(c) for c in cases]
        setup_s = time.perf_counter() - t0
        b = Batch(solvers)
        if params is not None:
            b.set_indicator(params, thr)
        t1 = time.perf_counter()
        threads, grid = b.tune(reps=2)
        tune_s = time.perf_counter() - t1
        b.counts(reset=True)
        t2 = time.perf_counter()
        dev_ms = b.march(steps, graph=True)
        wall_s = time.perf_counter() - t2
        lin, seen, over, flagged = b.counts()
        b.close()
        for s in solvers:
            s.close()
        for c in cases:
            c.close()
        # one case at a time, same steps, from the initial state
        per_case = []
        cases = [Case(d) for d in dirs]
        for c in cases:
            solver = Solver(c)
            if params is not None:
                solver.set_indicator(params, thr)
            solver.synchronize()
            ts = time.perf_counter()
            solver.march(steps)
            solver.synchronize()
            per_case.append((time.perf_counter() - ts) * 1e3)
            solver.close()
        for c in cases:
            c.close()
    total = sum(cells) * 18 * steps
    pc = np.array(per_case)
    sides = sorted(set(d["side"] for d in drawn))
    return {"cases": count, "steps": steps, "seed": seed, "indicator": params is not None,
            "value": round(total / wall_s / 1e6, 3), "unit": "Mcell-stage updates/s",
            "wall_ms": round(wall_s * 1e3, 3), "device_ms": round(dev_ms, 3),
            "serial_value": round(total / (pc.sum() / 1e3) / 1e6, 3),
            "serial_ms_per_case": {"p50": round(float(np.percentile(pc, 50)), 3),
                                   "p90": round(float(np.percentile(pc, 90)), 3),
                                   "p99": round(float(np.percentile(pc, 99)), 3), "sum": round(float(pc.sum()), 3)},
            "linear_face_fraction": round(lin / max(seen, 1), 4) if params is not None else None,
            "weno5_face_fraction": round(over / max(seen, 1), 4) if params is not None else None,
            "flagged_face_fraction": round(flagged / max(seen, 1), 4) if params is not None else None,
            "tuned": {"threads_per_line": {str(sd): sorted(set(t for t, d in zip(threads, drawn) if d["side"] == sd))
                                           for sd in sides}, "grid_blocks": grid},
            "setup_seconds": round(setup_s, 2), "tune_seconds": round(tune_s, 2),
            "cells": int(sum(cells))}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--case", default="shockvortex256")
    ap.add_argument("--segment", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-hybrid", action="store_true")
    ap.add_argument("--unfused", action="store_true", help="stage-by-stage sweeps instead of the fused line sweeps")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-case batched workload")
    ap.add_argument("--batched-only", action="store_true", help="run only the 64-case batched workload")
    ap.add_argument("--batched-plain", action="store_true", help="batched workload without the indicator run")
    ap.add_argument("--no-side", action="store_true", help="skip the viscous TGV and immersed-cylinder workloads")
    ap.add_argument("--side-only", action="store_true",
                    help="run only the viscous TGV, immersed-cylinder and moving-cylinder workloads")
    ap.add_argument("--no-end-time", action="store_true",
                    help="skip the plain and hybrid marches to the end time (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the multi-GPU slab decomposition is a "
                                   "design that is not built, so --gpus must be 1"}))
        sys.exit(2)
    from artracfd_amd import Case, Solver
    if args.batched_only:
        print(json.dumps({"batched": batched_both(not args.batched_plain)}))
        return
    if args.side_only:
        print(json.dumps({"viscous_and_immersed": side_workloads(args.warmup, not args.no_cpu)}))
        return
    case_dir = os.path.join(ROOT, "cases", args.case)
    t_init = time.perf_counter()
    case = Case(case_dir)
    cells = fluid_nodes(case)
    solver = Solver(case, segment=args.segment, fused=not args.unfused)
    fused_axes = solver.fused_axes()
    init_s = time.perf_counter() - t_init
    from artracfd_amd.lib import copy_probe
    copy_gbs = copy_probe(1 << 30, 20)
    done, elapsed, tm, _ = timed(solver, args.warmup, args.steps)
    solver.close()
    case.close()
    value = cells * 18 * done / elapsed / 1e6
    sweep_ms, sweep_n = tm["sweep"]
    stages_per_launch = 18.0 * done / max(sweep_n, 1)
    bytes_per_launch = cells * stage_bytes(1) * stages_per_launch
    avg_launch_s = sweep_ms / 1e3 / max(sweep_n, 1)
    achieved = bytes_per_launch / avg_launch_s / 1e9
    flops_per_launch = FP64_FLOP_PER_CELL_STAGE * cells * stages_per_launch
    traffic = None
    if args.case == "shockvortex256":
        per = TRAFFIC_B_PER_CELL_STAGE if not args.unfused else UNFUSED_TRAFFIC_B_PER_CELL_STAGE
        traffic = round(per * cells * stages_per_launch)
    kernel = ("rk_line (fused directional sweep, all RK stages of one line in LDS)" if fused_axes
              else "sweep_x/sweep_march (stage-by-stage directional sweeps)")
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
        "config": {"workload": describe(case_dir, args.case, cells), "parallelism": "single GPU",
                   "fused_axes": fused_axes},
        "roofline": {"bound": "hbm", "limiter": "FP64 vector issue (fp64 below); HBM traffic is a quarter of the "
                                                "algorithmic bytes", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kernel, "avg_launch_ms": round(avg_launch_s * 1e3, 4), "launches": sweep_n,
                     "cell_stages_per_launch": round(cells * stages_per_launch),
                     "algorithmic_bytes_per_launch": round(bytes_per_launch),
                     "measured_copy_gbs": round(copy_gbs, 1),
                     "fp64": {"achieved_tflops": round(flops_per_launch / avg_launch_s / 1e12, 2),
                              "peak_tflops": FP64_PEAK_TFLOPS,
                              "frac": round(flops_per_launch / avg_launch_s / 1e12 / FP64_PEAK_TFLOPS, 4),
                              "flop_per_cell_stage": FP64_FLOP_PER_CELL_STAGE}},
        "kernel_ms": {k: round(v[0], 3) for k, v in tm.items()},
        "kernel_launches": {k: v[1] for k, v in tm.items()},
        "init_seconds": round(init_s, 2),
    }
    from artracfd_amd import indicator
    if not args.no_hybrid and not args.unfused and not os.path.exists(indicator.WEIGHTS):
        out["hybrid"] = {"error": f"no trained indicator weights at {os.path.relpath(indicator.WEIGHTS, ROOT)}"}
    elif not args.no_hybrid and not args.unfused:
        params, meta = indicator.load()
        thr = float(meta.get("logit_threshold", indicator.LOGIT_THRESHOLD))
        case = Case(case_dir)
        solver = Solver(case)
        solver.set_indicator(params, thr)
        hdone, helapsed, htm, (lin, seen, over, flagged) = timed(solver, args.warmup, args.steps)
        solver.close()
        case.close()
        hvalue = cells * 18 * hdone / helapsed / 1e6
        if args.no_end_time:
            rho_p = rho_h = np.ones(1)
            steps_p = steps_h = el_p = el_h = 0
        else:
            rho_p, steps_p, el_p = run_to_end(case_dir)
            rho_h, steps_h, el_h = run_to_end(case_dir, params, thr)
        out["hybrid"] = {
            "value": round(hvalue, 3), "unit": "Mcell-stage updates/s", "vs_plain": round(hvalue / value, 3),
            "ms_per_step": round(helapsed * 1e3 / max(hdone, 1), 4),
            "linear_face_fraction": round(lin / max(seen, 1), 4),
            "weno5_face_fraction": round(over / max(seen, 1), 4),
            "flagged_face_fraction": round(flagged / max(seen, 1), 4),
            "rel_l1_rho_vs_plain_at_end": float(np.abs(rho_h - rho_p).sum() / np.abs(rho_p).sum()),
            "end_time_steps": {"plain": steps_p, "hybrid": steps_h},
            "end_time_seconds": {"plain": round(el_p, 3), "hybrid": round(el_h, 3)},
            "logit_threshold": thr, "kernel_ms": {k: round(v[0], 3) for k, v in htm.items()},
        }
    if not args.no_batched and not args.unfused:
        out["batched"] = batched_both(not args.batched_plain)
    if not args.no_side and not args.unfused:
        out["viscous_and_immersed"] = side_workloads(args.warmup, not args.no_cpu)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
