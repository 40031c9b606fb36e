"""CPU restatement of the learned indicator's network (numpy).

TEST INFRASTRUCTURE ONLY: the checker for the device training and inference
kernels (artracfd_amd/csrc/device/indicator.hip), used by tests/ only.

The device computes the 12-32-32-1 ReLU network with BF16 operands on MFMA
and FP32 accumulation.  This restatement rounds the same operands to BF16
(round to nearest even) at the same points - inputs, W1, W2, the layer-1
activations fed to layer 2 and to dW2, the back-propagated dZ2 and dZ1 fed
to the matrix products - and accumulates in FP64, so it agrees with the
device to FP32 accumulation-order rounding.

Parameter vector layout (1633 FP32 values): W1[16][32] (rows 12..15 zero),
b1[32], W2[32][32], b2[32], w3[32], b3.
"""
import numpy as np

P = 1633
O_W1, O_B1, O_W2, O_B2, O_W3, O_B3 = 0, 512, 544, 1568, 1600, 1632


def bf16(x):
    """Round FP32 values to BF16 (nearest even) and return them as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    W1 = p[O_W1:O_B1].reshape(16, 32)
    W2 = p[O_W2:O_B2].reshape(32, 32)
    return W1, p[O_B1:O_W2], W2, p[O_B2:O_W3], p[O_W3:O_B3], p[O_B3]


def init_params(seed=0):
    """He-initialised parameters (FP32), padded rows of W1 zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(P, dtype=np.float32)
    W1 = np.zeros((16, 32))
    W1[:12] = rng.normal(0, np.sqrt(2.0 / 12), (12, 32))
    p[O_W1:O_B1] = W1.ravel()
    p[O_W2:O_B2] = rng.normal(0, np.sqrt(2.0 / 32), (32, 32)).ravel()
    p[O_W3:O_B3] = rng.normal(0, np.sqrt(1.0 / 32), 32)
    return p


def _identity(x):
    return np.asarray(x, dtype=np.float64)


def forward(params, x, rounding=True):
    """Logits of feature rows x [n][12]; returns (z, cache).  rounding=False
    keeps every operand in FP64 (the smooth network, for finite differences)."""
    r = bf16 if rounding else _identity
    W1, b1, W2, b2, w3, b3 = unpack(params)
    X = np.zeros((len(x), 16))
    X[:, :12] = r(np.asarray(x, dtype=np.float32))
    H1 = np.maximum(X @ r(W1) + b1, 0.0)
    H1b = r(H1)
    H2 = np.maximum(H1b @ r(W2) + b2, 0.0)
    z = H2 @ w3 + b3
    return z, (X, H1, H1b, H2)


def loss_value(params, x, labels, tau, wpos, rounding=True):
    """Weighted logistic loss of (label > tau), weights normalised to sum 1."""
    z, _ = forward(params, x, rounding)
    y = (np.asarray(labels) > tau).astype(np.float64)
    w = np.where(y > 0.5, wpos, 1.0)
    w = w / w.sum()
    return float(np.sum(w * (np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))))


def gradient(params, x, labels, tau, wpos, rounding=True):
    """Weighted logistic loss of (label > tau) and its gradient."""
    r = bf16 if rounding else _identity
    W1, b1, W2, b2, w3, b3 = unpack(params)
    z, (X, H1, H1b, H2) = forward(params, x, rounding)
    y = (np.asarray(labels) > tau).astype(np.float64)
    w = np.where(y > 0.5, wpos, 1.0)
    w = w / w.sum()
    sig = 1.0 / (1.0 + np.exp(-z))
    dl = (sig - y) * w
    loss = np.sum(w * (np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))))
    g = np.zeros(P)
    g[O_W3:O_B3] = H2.T @ dl
    g[O_B3] = dl.sum()
    dZ2 = np.where(H2 > 0, dl[:, None] * w3[None, :], 0.0)
    g[O_B2:O_W3] = dZ2.sum(axis=0)
    dZ2b = r(dZ2)
    g[O_W2:O_B2] = (H1b.T @ dZ2b).ravel()
    dH1 = dZ2b @ r(W2).T
    dZ1 = np.where(H1b > 0, dH1, 0.0)
    g[O_B1:O_W2] = dZ1.sum(axis=0)
    dW1 = X.T @ r(dZ1)
    dW1[12:] = 0.0
    g[O_W1:O_B1] = dW1.ravel()
    return g, loss
