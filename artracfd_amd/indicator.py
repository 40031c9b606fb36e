"""Training data, training and evaluation of the learned troubled-face indicator.

Data come from the solver's own plain WENO5 runs on the device: seeded random
two-dimensional Riemann problems and sinusoidal density waves (collapsed z
axis, outflow or periodic boundaries).  Every few steps every face along x
and y is sampled (features and the WENO5 weight-deviation label), and a
seeded subsample is kept.  Training runs on the device kernels
(artra_mlp_train); the weights are one FP32 vector of 1633 values.

    python -m artracfd_amd.indicator train --samples 20000000 --out weights.npz
"""
import argparse
import os
import tempfile

import numpy as np

from . import lib

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "weights", "indicator.npz")

# label threshold of the committed weights and the routing threshold chosen
# on the 64^3 shock-vortex run
TAU = 0.4
LOGIT_THRESHOLD = 2.0


def _case_text(mx, my, end, steps, init, bcs, regions):
    lines = ["space begin", "0, 0, 0", "1, 1, 0.01", f"{mx}, {my}, 1", "space end",
             "time begin", "0", repr(end), "0.6", str(steps), "1", "0", "time end",
             "numerical begin", "1", "1", "0", "0", "0", "0", "1", "numerical end",
             "material begin", "0", "0.0", "0", "0, 0, 0", "material end",
             "reference begin", "1.0", "1.0", "1.0", "1.0", "reference end",
             "initialization begin"] + list(init) + ["initialization end"]
    for name, b in zip(("west", "east", "south", "north", "front", "back"), bcs):
        lines += [f"{name} boundary begin", b, f"{name} boundary end"]
    for lo, hi, st in regions:
        lines += ["box initialization begin", f"{lo[0]}, {lo[1]}, -1", f"{hi[0]}, {hi[1]}, 1"]
        lines += [f"{v:.6f}" for v in st] + ["box initialization end"]
    return "\n".join(lines) + "\n"


def _write(d, text):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "artracfd.case"), "w") as fp:
        fp.write(text)
    with open(os.path.join(d, "artracfd.geo"), "w") as fp:
        fp.write("count begin\n0\n0\ncount end\n")
    return d


def riemann_case(d, rng, n=96, steps=120):
    """Four-quadrant Riemann problem with random states, outflow boundaries."""
    xc, yc = rng.uniform(0.3, 0.7, 2)
    st = [(rng.uniform(0.1, 2.0), rng.uniform(-1, 1), rng.uniform(-1, 1), 0.0, rng.uniform(0.1, 2.0))
          for _ in range(4)]
    init = [f"{st[0][0]:.6f}", f"{st[0][1]:.6f}", f"{st[0][2]:.6f}", "0", f"{st[0][4]:.6f}"]
    regions = [((-1, -1), (xc, yc), st[1]), ((xc, -1), (2, yc), st[2]), ((-1, yc), (xc, 2), st[3])]
    return _write(d, _case_text(n, n, 1e9, steps, init, ["outflow"] * 6, regions))


def wave_case(d, n=96, steps=120, kx=1, ky=1, amplitude=0.2, u=0.5, v=0.25):
    """Density wave advected through a periodic box at uniform pressure 1."""
    init = [f"1+{amplitude:.6g}*sin(2*pi*({kx}*x+{ky}*y))", f"{u:.5f}", f"{v:.5f}", "0", "1"]
    return _write(d, _case_text(n, n, 1e9, steps, init, ["periodic"] * 4 + ["outflow"] * 2, []))


def training_cases(root, seed, count=32, n=96, steps=120):
    """Write `count` case directories under root (alternately Riemann
    problems and density waves); returns their paths."""
    rng = np.random.default_rng(seed)
    dirs = []
    for c in range(count):
        d = os.path.join(root, f"train{c:03d}")
        if c % 2 == 0:
            dirs.append(riemann_case(d, rng, n, steps))
        else:
            kx, ky = rng.integers(1, 6, 2)
            dirs.append(wave_case(d, n, steps, int(kx), int(ky), rng.uniform(0.05, 0.6), *rng.uniform(-1, 1, 2)))
    return dirs


def collect(samples, seed=1, every=3, keep=0.35, count=32, n=96, steps=120, verbose=False):
    """Features [m][12] and labels [m] from plain WENO5 runs (m <= samples)."""
    rng = np.random.default_rng(seed + 17)
    feats, labs, total = [], [], 0
    with tempfile.TemporaryDirectory() as root:
        rounds = 0
        while total < samples:
            for d in training_cases(root, seed + 1000 * rounds, count, n, steps):
                case = lib.Case(d)
                solver = lib.Solver(case)
                for _ in range(steps // every):
                    done, _ = solver.march(every)
                    if done == 0:
                        break
                    ok = True
                    for axis in (0, 1):
                        f, y = solver.sample_faces(axis)
                        finite = np.isfinite(y) & np.all(np.isfinite(f), axis=1)
                        ok = ok and bool(finite.all())
                        pick = (rng.random(len(y)) < keep) & finite
                        feats.append(f[pick])
                        labs.append(y[pick])
                        total += int(pick.sum())
                    if not ok:
                        break
                solver.close()
                case.close()
                if total >= samples:
                    break
            rounds += 1
            if verbose:
                print(f"collected {total} samples", flush=True)
    F = np.concatenate(feats)[:samples]
    Y = np.concatenate(labs)[:samples]
    order = np.random.default_rng(seed + 99).permutation(len(Y))
    return np.ascontiguousarray(F[order]), np.ascontiguousarray(Y[order])


def train(features, labels, seed=0, epochs=4, batch=4096, lr=3e-3, tau=TAU, wpos=4.0):
    p0 = _init_params(seed)
    n = (len(labels) // batch) * batch
    return lib.mlp_train(p0, features[:n], labels[:n], epochs=epochs, batch=batch, lr=lr, tau=tau, wpos=wpos)


def _init_params(seed):
    """He-initialised parameters (the same recipe as the oracle's init_params)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(lib.MLP_PARAMS, dtype=np.float32)
    W1 = np.zeros((16, 32))
    W1[:12] = rng.normal(0, np.sqrt(2.0 / 12), (12, 32))
    p[0:512] = W1.ravel()
    p[544:1568] = rng.normal(0, np.sqrt(2.0 / 32), (32, 32)).ravel()
    p[1600:1632] = rng.normal(0, np.sqrt(1.0 / 32), 32)
    return p


def save(path, params, **meta):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=np.asarray(params, dtype=np.float32), **meta)


def load(path=WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32), {k: z[k] for k in z.files if k != "params"}


def classification_report(params, features, labels, tau=TAU, logit_threshold=LOGIT_THRESHOLD):
    z = lib.mlp_logits(params, features)
    pred = z > logit_threshold
    truth = labels > tau
    return {"samples": int(len(labels)), "troubled_fraction": float(truth.mean()),
            "routed_linear_fraction": float((~pred).mean()),
            "missed_troubled": float((truth & ~pred).sum() / max(truth.sum(), 1)),
            "false_troubled": float((~truth & pred).sum() / max((~truth).sum(), 1))}


def evaluate(case_directory, params, thresholds, max_steps=1 << 20):
    """Plain WENO5 and routed runs of one case to its end time: per logit
    threshold the relative L1 density difference, the linear fraction of
    the faces and both run times (s)."""
    import time

    def run(p, thr):
        case = lib.Case(case_directory)
        solver = lib.Solver(case)
        if p is not None:
            solver.set_indicator(p, thr)
        solver.hybrid_counts(reset=True)
        solver.synchronize()
        t0 = time.perf_counter()
        steps, _ = solver.march(max_steps)
        solver.synchronize()
        el = time.perf_counter() - t0
        rho = case.interior(solver.field())[0]
        counts = solver.hybrid_counts()
        solver.close()
        case.close()
        return rho, steps, el, counts

    rho0, steps0, el0, _ = run(None, 0.0)
    out = []
    for thr in thresholds:
        rho, steps, el, (lin, seen, _, _) = run(params, thr)
        out.append({"threshold": float(thr), "rel_l1_rho": float(np.abs(rho - rho0).sum() / np.abs(rho0).sum()),
                    "linear_fraction": lin / max(seen, 1), "steps": steps, "plain_steps": steps0,
                    "seconds": el, "plain_seconds": el0})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("command", choices=["train"])
    ap.add_argument("--samples", type=int, default=20_000_000)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--out", default=WEIGHTS)
    args = ap.parse_args()
    import time
    t0 = time.time()
    F, Y = collect(args.samples, seed=args.seed, verbose=True)
    t1 = time.time()
    params, hist = train(F, Y, seed=args.seed, epochs=args.epochs)
    t2 = time.time()
    rep = classification_report(params, F[:1_000_000], Y[:1_000_000])
    print({"collect_s": round(t1 - t0, 1), "train_s": round(t2 - t1, 1), "loss": hist.tolist(), **rep}, flush=True)
    save(args.out, params, tau=TAU, logit_threshold=LOGIT_THRESHOLD, loss=hist, samples=len(Y), seed=args.seed,
         label_quantiles=np.quantile(Y, [0.5, 0.9, 0.99]))


if __name__ == "__main__":
    main()
