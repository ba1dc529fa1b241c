"""Test infrastructure only: float64 numpy restatement of the proposal MLP
(256 -> 256 -> 256 -> 1, ReLU, mean squared error) used to check the GPU
forward pass and gradients.

With ``bf16=True`` inputs, weights and the operands of every matrix product
are rounded to bfloat16 (round to nearest even) at the points where the GPU
rounds them, and everything else is float64; the GPU accumulates in fp32, so
the two agree to fp32 accumulation error.  With ``bf16=False`` it is the plain
float64 model, whose gradients are pinned by central finite differences in
tests/test_mlp_oracle.py.
"""
from __future__ import annotations

import numpy as np

IN = HID = 256
SIZES = [("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
         ("w3", (HID,)), ("b3", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def bf16_round(x) -> np.ndarray:
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d) -> np.ndarray:
    return np.concatenate([np.asarray(d[n], np.float64).reshape(-1) for n, _ in SIZES])


def forward(p, X, bf16=True):
    P = unpack(p)
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    Xb = r(X)
    Z1 = Xb @ r(P["W1"]) + P["b1"]
    A1 = r(np.maximum(Z1, 0))
    Z2 = A1 @ r(P["W2"]) + P["b2"]
    A2 = r(np.maximum(Z2, 0))
    y = A2 @ r(P["w3"]) + P["b3"][0]
    return y, dict(Xb=Xb, Z1=Z1, A1=A1, Z2=Z2, A2=A2)


def loss_and_grads(p, X, t, bf16=True):
    P = unpack(p)
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    y, c = forward(p, X, bf16)
    B = X.shape[0]
    e = y - np.asarray(t, np.float64)
    loss = float(np.mean(e * e))
    dy = 2.0 * e / B
    g = {}
    g["w3"] = c["A2"].T @ dy
    g["b3"] = np.array([dy.sum()])
    dZ2 = (dy[:, None] * r(P["w3"])[None, :]) * (c["Z2"] > 0)
    dZ2b = r(dZ2)
    g["W2"] = c["A1"].T @ dZ2b
    g["b2"] = dZ2.sum(0)
    dZ1 = (dZ2b @ r(P["W2"]).T) * (c["Z1"] > 0)
    dZ1b = r(dZ1)
    g["W1"] = c["Xb"].T @ dZ1b
    g["b1"] = dZ1.sum(0)
    return loss, pack(g)


def init_params(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    d = {"W1": rng.uniform(-1, 1, (IN, HID)) * np.sqrt(6 / IN),
         "b1": rng.uniform(-0.1, 0.1, HID),
         "W2": rng.uniform(-1, 1, (HID, HID)) * np.sqrt(6 / HID),
         "b2": rng.uniform(-0.1, 0.1, HID),
         "w3": rng.uniform(-1, 1, HID) * np.sqrt(6 / HID),
         "b3": np.array([0.05])}
    return pack(d)
