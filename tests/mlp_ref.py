"""CPU FP64 reference of the surrogate network (test infrastructure): the
same parameter layout and loss as noa_amd/csrc/mlp_common.hh and
surrogate_train.hip, in numpy float64. With `bf16=True` the operands of the
MFMA layers (activations, weights and back-propagated gradients, and both
operands of the hidden weight gradients) are rounded to bf16 exactly where
the GPU rounds them (the biases, W1 and W5 gradients are FP32 there), so the
GPU's gradients can be compared tightly; with `bf16=False` it is the exact
network."""
from __future__ import annotations

import numpy as np

IN, H, OUT = 3, 128, 4


def layout():
    off, o = {}, 0
    for name, size in (("W1", H * IN), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H * H), ("b3", H),
                       ("W4", H * H), ("b4", H), ("W5", OUT * H), ("b5", OUT)):
        off[name] = (o, size)
        o += size
    return off, o


OFF, NPARAM = layout()


def unpack(p: np.ndarray) -> dict:
    out = {}
    for name, (o, s) in OFF.items():
        v = p[o:o + s].astype(np.float64)
        if name[0] == "W":
            v = v.reshape(-1, IN if name == "W1" else H)
        out[name] = v
    return out


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) and back."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p: np.ndarray, X: np.ndarray, round_bf16: bool = False):
    q = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    w = unpack(p)
    Z, A = [], []
    z = X.astype(np.float64) @ w["W1"].T + w["b1"]
    Z.append(z)
    A.append(z * sigmoid(z))
    for l in (2, 3, 4):
        z = q(A[-1]) @ q(w[f"W{l}"]).T + w[f"b{l}"]
        Z.append(z)
        A.append(z * sigmoid(z))
    Y = q(A[-1]) @ q(w["W5"]).T + w["b5"]
    return Y, Z, A


def loss_and_grad(p: np.ndarray, X, T, M, round_bf16: bool = False):
    q = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    w = unpack(p)
    Y, Z, A = forward(p, X, round_bf16)
    n = X.shape[0]
    d = (Y - T) * M
    loss = float(np.sum(d * d) / (n * OUT))
    dY = 2.0 * d / (n * OUT)
    g = {}
    g["W5"] = dY.T @ A[3]  # FP32 on the GPU
    g["b5"] = dY.sum(0)
    dA = q(dY) @ q(w["W5"])
    for l, zi in ((4, 3), (3, 2), (2, 1)):
        s = sigmoid(Z[zi])
        dZ = dA * s * (1 + Z[zi] * (1 - s))
        g[f"W{l}"] = q(dZ).T @ q(A[zi - 1])
        g[f"b{l}"] = dZ.sum(0)
        dA = q(dZ) @ q(w[f"W{l}"])
    s = sigmoid(Z[0])
    dZ = dA * s * (1 + Z[0] * (1 - s))
    g["W1"] = dZ.T @ X.astype(np.float64)
    g["b1"] = dZ.sum(0)
    flat = np.zeros(NPARAM)
    for name, (o, size) in OFF.items():
        flat[o:o + size] = g[name].reshape(-1)
    return loss, flat
