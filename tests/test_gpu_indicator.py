"""The learned troubled-face indicator on the device: the BF16-MFMA network
against the CPU restatement (forward and gradient), training, the sampled
features and labels, and the routed sweeps against plain WENO5."""
import os

import numpy as np
import pytest

import mlp_oracle as M
from conftest import case_dir

pytestmark = pytest.mark.gpu


def _params(seed):
    rng = np.random.default_rng(seed)
    p = M.init_params(seed)
    p[M.O_B1:M.O_W2] = rng.normal(0, 0.1, 32)
    p[M.O_B2:M.O_W3] = rng.normal(0, 0.1, 32)
    p[M.O_B3] = -0.2
    return p


@pytest.mark.parametrize("n", [64, 4096 + 37])
def test_logits_match_oracle(n):
    """Same BF16 operands, FP32 (device) vs FP64 (oracle) accumulation: a
    rare BF16 tie flip of a hidden value is the only larger difference."""
    from artracfd_amd.lib import mlp_logits
    print(f"logits n={n}", flush=True)
    rng = np.random.default_rng(n)
    p = _params(1)
    x = rng.uniform(-1, 1, (n, 12)).astype(np.float32)
    z = mlp_logits(p, x).astype(np.float64)
    zo, _ = M.forward(p, x)
    d = np.abs(z - zo)
    assert np.median(d) <= 1e-5 * (1 + np.median(np.abs(zo)))
    assert np.quantile(d, 0.999) <= 2e-3 * (1 + np.abs(zo).max())
    assert d.max() <= 2e-2 * (1 + np.abs(zo).max())


@pytest.mark.parametrize("n", [1000, 8192])
def test_gradient_matches_oracle(n):
    """Weighted logistic loss and its gradient (forward and backward on
    MFMA) against the oracle's back-propagation, which finite differences
    pin on the CPU."""
    from artracfd_amd.lib import mlp_gradient
    print(f"gradient n={n}", flush=True)
    rng = np.random.default_rng(10 + n)
    p = _params(2)
    x = rng.uniform(-1, 1, (n, 12)).astype(np.float32)
    y = rng.random(n).astype(np.float32)
    g, loss = mlp_gradient(p, x, y, 0.3, 3.0)
    go, lo = M.gradient(p, x, y.astype(np.float64), 0.3, 3.0)
    assert loss == pytest.approx(lo, rel=1e-5)
    assert np.linalg.norm(g - go) <= 2e-3 * np.linalg.norm(go)
    for a, b in ((M.O_W1, M.O_B1), (M.O_B1, M.O_W2), (M.O_W2, M.O_B2), (M.O_B2, M.O_W3), (M.O_W3, M.O_B3)):
        assert np.linalg.norm(g[a:b] - go[a:b]) <= 1e-2 * np.linalg.norm(go[a:b]) + 1e-9, (a, b)
    assert g[M.O_B3] == pytest.approx(go[M.O_B3], rel=1e-4, abs=1e-8)
    assert np.all(g[12 * 32:M.O_B1] == 0.0)


def test_training_learns_and_is_deterministic():
    """A separable rule on the features: the loss falls well below its
    starting value, and two runs give bit-identical parameters."""
    from artracfd_amd.lib import mlp_train
    print("training", flush=True)
    rng = np.random.default_rng(4)
    n = 1 << 16
    x = rng.uniform(-1, 1, (n, 12)).astype(np.float32)
    y = (np.abs(x[:, 2] - x[:, 3]) + 0.5 * np.abs(x[:, 8] - x[:, 9])).astype(np.float32)
    p0 = M.init_params(4)
    l0 = M.loss_value(p0, x, y, 0.8, 2.0)
    p1, hist = mlp_train(p0, x, y, epochs=6, batch=1024, lr=3e-3, tau=0.8, wpos=2.0)
    p2, _ = mlp_train(p0, x, y, epochs=6, batch=1024, lr=3e-3, tau=0.8, wpos=2.0)
    assert np.array_equal(p1, p2)
    l1 = M.loss_value(p1, x, y, 0.8, 2.0)
    z, _ = M.forward(p1, x)
    acc = np.mean((z > 0) == (y > 0.8))
    print("loss", l0, "->", l1, "history", hist, "accuracy", acc)
    assert l1 < 0.25 * l0 and acc > 0.9
    assert np.all(p1[12 * 32:M.O_B1] == 0.0)


def _sod():
    from artracfd_amd import Case, Solver
    case = Case(case_dir("sod"))
    return case, Solver(case)


def test_sampled_features_are_the_normalised_stencils():
    """Features of every x face of the Sod tube after a few steps, against
    the same formula evaluated in numpy (FP32) on the downloaded field."""
    print("features sod", flush=True)
    case, solver = _sod()
    solver.march(5)
    U = solver.field()
    feat, lab = solver.sample_faces(0)
    (x0, x1), (y0, y1), (z0, z1) = case.pin
    faces = x1 - x0 + 1
    assert feat.shape == ((z1 - z0) * (y1 - y0) * faces, 12)
    f = feat.reshape(z1 - z0, y1 - y0, faces, 12)
    g1 = np.float32(0.4)
    V = U.astype(np.float32)
    rho = V[0]
    p = g1 * (V[4] - np.float32(0.5) * ((V[1] * V[1] + V[2] * V[2]) + V[3] * V[3]) * (np.float32(1) / V[0]))
    for q in range(faces):
        base = x0 - 1 + q
        for v, arr in enumerate((rho, p)):
            s = arr[z0:z1, y0:y1, base - 2:base + 4]
            c = np.float32(0.5) * (s[..., 2] + s[..., 3])
            dev = np.max(np.abs(s - c[..., None]), axis=-1)
            want = (s - c[..., None]) * (np.float32(1) / (dev + np.float32(0.01) * np.abs(c) + np.float32(1e-30)))[..., None]
            np.testing.assert_allclose(f[:, :, q, 6 * v:6 * v + 6], want, rtol=0, atol=2e-6)
    assert np.all(np.isfinite(lab)) and np.all(lab >= 0)
    solver.close()
    case.close()


def test_labels_mark_the_discontinuity():
    """On the initial Sod state, faces whose six-node stencil straddles the
    jump have WENO weights far from linear; all others have exactly linear
    weights (label zero to rounding).  A jump at the outer end of a
    five-node side stencil zeroes one candidate only: weights (0, 2/3, 1/3)
    against (0.1, 0.6, 0.3), a deviation of exactly 0.2; nearer the face
    the deviation is larger."""
    print("labels sod", flush=True)
    case, solver = _sod()
    feat, lab = solver.sample_faces(0)
    (x0, x1), (y0, y1), (z0, z1) = case.pin
    faces = x1 - x0 + 1
    U = solver.field()
    rho_line = U[0, z0, y0]
    jump = int(np.nonzero(np.diff(rho_line) != 0)[0][0])     # rho[jump] != rho[jump+1]
    L = lab.reshape(-1, faces)
    for q in range(faces):
        base = x0 - 1 + q
        straddles = base - 2 <= jump < base + 3
        if straddles:
            inner = base - 1 <= jump < base + 2
            assert np.all(L[:, q] > (0.5 if inner else 0.2 - 1e-6)), (q, L[:, q].min())
        else:
            assert np.all(L[:, q] < 1e-10), (q, L[:, q].max())
    solver.close()
    case.close()


def _evolve(name_or_dir, steps, params=None, threshold=0.0, gate=1e-3):
    from artracfd_amd import Case, Solver
    case = Case(name_or_dir if os.path.isdir(name_or_dir) else case_dir(name_or_dir))
    solver = Solver(case)
    if params is not None:
        solver.set_indicator(params, threshold, smooth_variation=gate)
    solver.hybrid_counts(reset=True)
    done, dts = solver.march(steps)
    U = case.interior(solver.field())
    counts = solver.hybrid_counts()
    solver.close()
    case.close()
    return U, dts, counts


def _flux_model_variant(tmp_path, model):
    """shockvortex64 with the Roe average (numerical block line 4) or the
    Steger-Warming split (line 5) in place of the default mean state and
    Lax-Friedrichs split."""
    if model == "llf":
        return case_dir("shockvortex64")
    src = case_dir("shockvortex64")
    dst = tmp_path / f"shockvortex64_{model}"
    dst.mkdir()
    lines = open(os.path.join(src, "artracfd.case")).read().split("\n")
    at = lines.index("numerical begin") + (4 if model == "roe" else 5)
    assert lines[at].strip() == "0"
    lines[at] = "1"
    (dst / "artracfd.case").write_text("\n".join(lines))
    (dst / "artracfd.geo").write_text(open(os.path.join(src, "artracfd.geo")).read())
    return str(dst)


@pytest.mark.parametrize("model", ["llf", "roe", "sw"])
def test_all_troubled_routing_is_plain_weno5(model, tmp_path):
    """With every face routed to WENO5 the hybrid sweeps reproduce the plain
    sweeps (same face state, projection and update arithmetic), for the
    Lax-Friedrichs split on the mean state and for the Roe average and the
    Steger-Warming split, which run a kernel variant of their own."""
    name = _flux_model_variant(tmp_path, model)
    print("hybrid all-WENO shockvortex64", model, flush=True)
    U0, d0, _ = _evolve(name, 4)
    U1, d1, (lin, seen, weno, _) = _evolve(name, 4, _params(3), threshold=-1e30, gate=0.0)
    assert lin == 0 and seen > 0
    assert weno == seen
    np.testing.assert_array_equal(d1, d0)
    assert np.array_equal(U1, U0)


def test_all_troubled_short_lines_are_plain_weno5(tmp_path):
    """Lines of 12 interior nodes (13 faces, one partly filled wavefront per
    line) with every face troubled: a 2-D Riemann problem matches the plain
    sweeps bit for bit."""
    from artracfd_amd import indicator
    d = indicator.riemann_case(str(tmp_path / "riemann12"), np.random.default_rng(7), n=12, steps=30)
    print("hybrid all-WENO on 13-face lines, riemann 12x12", flush=True)
    U0, d0, _ = _evolve(d, 30)
    U1, d1, (lin, seen, weno, _) = _evolve(d, 30, _params(3), threshold=-1e30, gate=0.0)
    assert lin == 0 and seen > 0 and weno == seen
    np.testing.assert_array_equal(d1, d0)
    assert np.array_equal(U1, U0)


def test_all_linear_on_smooth_data_is_close_to_weno5(tmp_path):
    """A small-amplitude density wave: every face routed to the linear
    fifth-order flux must agree with WENO5 (whose weights are then nearly
    linear) far below the size of the change over the run."""
    from artracfd_amd import Case, indicator
    print("hybrid all-linear density wave", flush=True)
    d = indicator.wave_case(str(tmp_path / "wave"), n=64, steps=20, kx=2, ky=1, amplitude=1e-3, u=0.5, v=0.25)
    init = Case(d)
    Ui = init.interior(init.field())
    init.close()
    Uw, dw, _ = _evolve(d, 20)
    Ul, dl, (lin, seen, weno, _) = _evolve(d, 20, _params(3), threshold=1e30)
    assert lin == seen and seen > 0 and weno == 0
    np.testing.assert_allclose(dl, dw, rtol=1e-8)    # the speeds differ at the level the fluxes do
    change = np.abs(Uw[0] - Ui[0]).max()
    assert change > 1e-5
    assert np.abs(Ul[0] - Uw[0]).max() <= 1e-3 * change


def test_trained_indicator_on_shock_vortex():
    """Committed weights on the 64^3 shock-vortex run to t = 0.6: density
    within 1e-3 relative L1 of plain WENO5 while most faces take the linear
    flux."""
    from artracfd_amd import indicator
    print("trained indicator shockvortex64", flush=True)
    params, meta = indicator.load()
    thr = float(meta.get("logit_threshold", indicator.LOGIT_THRESHOLD))
    Up, dp, _ = _evolve("shockvortex64", 1 << 20)
    Uh, dh, (lin, seen, over, _) = _evolve("shockvortex64", 1 << 20, params, thr)
    l1 = np.abs(Uh[0] - Up[0]).sum() / np.abs(Up[0]).sum()
    print("steps", len(dp), len(dh), "rel L1 rho", l1, "linear fraction", lin / seen)
    assert l1 <= 1e-3
    assert lin / seen >= 0.5


@pytest.mark.parametrize("model", ["roe", "sw"])
def test_trained_indicator_with_roe_and_steger_warming(model, tmp_path):
    """The routed path of the Roe-average and Steger-Warming flux models
    (linear flux through the characteristic projection) on the 64^3
    shock-vortex case for 20 steps: density within 1e-3 relative L1 of
    plain WENO5, mostly linear faces."""
    from artracfd_amd import indicator
    name = _flux_model_variant(tmp_path, model)
    print("trained indicator shockvortex64", model, flush=True)
    params, meta = indicator.load()
    thr = float(meta.get("logit_threshold", indicator.LOGIT_THRESHOLD))
    Up, dp, _ = _evolve(name, 20)
    Uh, dh, (lin, seen, over, _) = _evolve(name, 20, params, thr)
    l1 = np.abs(Uh[0] - Up[0]).sum() / np.abs(Up[0]).sum()
    print("rel L1 rho", l1, "linear fraction", lin / seen)
    assert l1 <= 1e-3
    assert lin / seen >= 0.5


def test_trained_network_gains_over_routing_without_it():
    """What the trained network adds, on the 64^3 shock-vortex run to
    t = 0.6: without a network the choice is between the smoothness gate
    alone (near-uniform faces linear, every other face WENO5) and the linear
    flux everywhere.  The network sends far more faces to the linear flux
    than the gate while staying within 1e-3 relative L1 of plain WENO5; the
    linear flux everywhere does not stay within it.  The shares are of the
    faces each choice flags troubled (the flux is then routed per wavefront
    of 64 faces, WENO5 where any face is flagged)."""
    from artracfd_amd import indicator
    params, meta = indicator.load()
    thr = float(meta.get("logit_threshold", indicator.LOGIT_THRESHOLD))

    def l1(U, Up):
        if not np.all(np.isfinite(U[0])):
            return np.inf
        return float(np.abs(U[0] - Up[0]).sum() / np.abs(Up[0]).sum())

    print("gain over routing without the network, shockvortex64", flush=True)
    Up, _, _ = _evolve("shockvortex64", 1 << 20)
    # per-face decisions: the share of faces not flagged troubled (the flux
    # itself is routed per wavefront of 64 faces)
    Un, _, (_, sn, _, fn) = _evolve("shockvortex64", 1 << 20, params, thr)
    Ug, _, (_, sg, _, fg) = _evolve("shockvortex64", 1 << 20, params, -1e30)      # gate only
    Ua, _, (_, sa, _, fa) = _evolve("shockvortex64", 1 << 20, params, 1e30, gate=0.0)  # all linear
    net, gate, lin = (l1(Un, Up), 1 - fn / sn), (l1(Ug, Up), 1 - fg / sg), (l1(Ua, Up), 1 - fa / sa)
    print("network", net, "gate only", gate, "all linear", lin)
    assert net[0] <= 1e-3
    assert lin[1] == 1.0 and lin[0] > 1e-3
    assert gate[0] <= 1e-3
    assert net[1] >= gate[1] + 0.1


def _ramp_case(tmp_path, numerical):
    """A density ramp rho = 1 + 0.5 x advected at u = 0.3 with uniform
    pressure, 128 cells along x (outflow ends), 4 x 4 periodic cells across:
    away from the ends every stencil along x is a linear ramp, so WENO5's
    smoothness indicators are all equal and its weights linear."""
    d = tmp_path / ("ramp_" + "".join(map(str, numerical)))
    d.mkdir()
    lines = ["space begin", "0, 0, 0", "1, 0.03125, 0.03125", "128, 4, 4", "space end",
             "time begin", "0", "1000000000.0", "0.6", "10", "1", "0", "time end",
             "numerical begin", *map(str, numerical), "numerical end",
             "material begin", "0", "0.0", "0", "0, 0, 0", "material end",
             "reference begin", "1.0", "1.0", "1.0", "1.0", "reference end",
             "initialization begin", "1+0.5*x", "0.3", "0", "0", "1", "initialization end"]
    for name, kind in zip(("west", "east", "south", "north", "front", "back"),
                          ("outflow", "outflow", "periodic", "periodic", "periodic", "periodic")):
        lines += [f"{name} boundary begin", kind, f"{name} boundary end"]
    (d / "artracfd.case").write_text("\n".join(lines) + "\n")
    (d / "artracfd.geo").write_text("count begin\n0\n0\ncount end\n")
    return str(d)


@pytest.mark.parametrize("numerical", [(1, 1, 0, 0, 0, 0, 1), (1, 1, 0, 1, 0, 0, 1), (1, 1, 0, 0, 1, 0, 1)],
                         ids=["llf", "roe", "sw"])
def test_all_linear_routing_equals_weno5_on_linear_data(tmp_path, numerical):
    """Every face routed to the linear flux (threshold never reached) against
    plain WENO5 on linear data, one time step: away from the outflow ends
    (which reach at most 18 nodes in: two x operators of three stages, three
    nodes each) the two agree to 1e-13 of each variable group's magnitude,
    for the mean-state and Roe-average Lax-Friedrichs and the Steger-Warming
    splits.  Pins the routed kernel's linear flux (stencil coefficients,
    face state from the per-node quantities, eigenvalue split)."""
    d = _ramp_case(tmp_path, numerical)
    print("all-linear vs WENO5 on a density ramp", numerical, flush=True)
    from artracfd_amd import Case
    c = Case(d)
    U0 = c.interior(c.field())
    c.close()
    Uw, dw, _ = _evolve(d, 1)
    Ul, dl, (lin, seen, weno, _) = _evolve(d, 1, _params(3), threshold=1e30, gate=0.0)
    assert lin == seen and seen > 0 and weno == 0
    mid = slice(24, -24)
    change = np.abs(Uw[0, ..., mid] - U0[0, ..., mid]).max()
    assert change > 1e-4, change
    for sl in (slice(0, 1), slice(1, 4), slice(4, 5)):
        scale = np.abs(Uw[sl, ..., mid]).max()
        diff = np.abs(Ul[sl, ..., mid] - Uw[sl, ..., mid]).max()
        print(sl, diff / scale)
        assert diff <= 1e-13 * scale, (sl, diff, scale)
    np.testing.assert_allclose(dl, dw, rtol=1e-14)


def test_trained_indicator_at_256_to_end_time():
    """The learned-hybrid workload itself: the 256^3 shock-vortex case to
    t = 0.6 with the committed weights against plain WENO5 on the same GPU:
    the same number of steps, density within 1e-3 relative L1, and most faces
    on the linear flux."""
    from artracfd_amd import indicator
    print("trained indicator shockvortex256 to t = 0.6", flush=True)
    params, meta = indicator.load()
    thr = float(meta.get("logit_threshold", indicator.LOGIT_THRESHOLD))
    rp, dp, _ = _evolve_density("shockvortex256")
    rh, dh, (lin, seen, over, flagged) = _evolve_density("shockvortex256", params, thr)
    l1 = float(np.abs(rh - rp).sum() / np.abs(rp).sum())
    print("steps", len(dp), len(dh), "rel L1 rho", l1, "linear", lin / seen, "weno5", over / seen,
          "flagged", flagged / seen, flush=True)
    assert len(dh) == len(dp) and len(dp) > 100
    assert l1 <= 1e-3
    # routed per wavefront of 64 faces: most wavefronts linear, few faces flagged
    assert lin / seen >= 0.8
    assert flagged / seen <= 0.05


def _evolve_density(name, params=None, threshold=0.0):
    """Density at the case's end time, the dt sequence and the face tallies
    (the full field of a 256^3 case is not kept)."""
    from artracfd_amd import Case, Solver
    case = Case(case_dir(name))
    solver = Solver(case)
    if params is not None:
        solver.set_indicator(params, threshold)
    solver.hybrid_counts(reset=True)
    done, dts = solver.march()
    rho = case.interior(solver.field())[0].copy()
    counts = solver.hybrid_counts()
    solver.close()
    case.close()
    return rho, dts, counts
