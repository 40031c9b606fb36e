"""The CPU restatement of the indicator network (oracle/mlp_oracle.py),
pinned by finite differences of its own loss and by torch's BF16 rounding;
and the host-side helpers of the indicator package."""
import numpy as np
import pytest

import mlp_oracle as M


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=20000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 20000),
                        np.float32([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, -0.0, 65504.0])])
    ref = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.bf16(x), ref)
    # ties go to even
    assert M.bf16(np.float32([1.0 + 2 ** -8]))[0] == 1.0
    assert M.bf16(np.float32([1.0 + 3 * 2 ** -8]))[0] == 1.0 + 2 ** -6


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_gradient_is_the_derivative_of_its_loss(seed):
    """Central differences of the unrounded (FP64) loss against the
    back-propagated gradient; a dropped term, a transposed weight or a wrong
    ReLU mask shows up as an O(1) relative error."""
    rng = np.random.default_rng(seed)
    p = M.init_params(seed).astype(np.float64)
    p[M.O_B1:M.O_W2] = rng.normal(0, 0.1, 32)
    p[M.O_B2:M.O_W3] = rng.normal(0, 0.1, 32)
    p[M.O_B3] = 0.3
    x = rng.uniform(-1, 1, (300, 12))
    y = rng.random(300)
    g, loss = M.gradient(p, x, y, 0.4, 3.0, rounding=False)
    assert loss == pytest.approx(M.loss_value(p, x, y, 0.4, 3.0, rounding=False), rel=1e-14)
    h = 1e-6
    fd = np.zeros_like(g)
    for q in range(M.P):
        if 12 * 32 <= q < M.O_B1:
            continue
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        fd[q] = (M.loss_value(pp, x, y, 0.4, 3.0, False) - M.loss_value(pm, x, y, 0.4, 3.0, False)) / (2 * h)
    assert np.linalg.norm(fd - g) <= 1e-7 * np.linalg.norm(g)
    assert np.all(g[12 * 32:M.O_B1] == 0.0)


def test_oracle_rounding_changes_little():
    """The BF16-rounded network stays within BF16 precision of the FP64 one."""
    rng = np.random.default_rng(3)
    p = M.init_params(3)
    x = rng.uniform(-1, 1, (2000, 12))
    z1, _ = M.forward(p, x)
    z0, _ = M.forward(p, x, rounding=False)
    assert np.max(np.abs(z1 - z0)) <= 0.05 * (1.0 + np.max(np.abs(z0)))
    assert np.max(np.abs(z1 - z0)) > 0.0


def test_training_cases_are_readable(tmp_path):
    from artracfd_amd import Case
    from artracfd_amd import indicator
    dirs = indicator.training_cases(str(tmp_path), seed=5, count=4, n=24, steps=6)
    for d in dirs:
        case = Case(d)
        U = case.interior(case.field())
        assert U.shape[0] == 5 and U.shape[1] == 1
        assert np.all(U[0] > 0) and np.all(np.isfinite(U))
        case.close()
