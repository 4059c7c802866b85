"""CPU restatement of the learned MAC's network (test infrastructure: only
tests/ use it, as the checker of the GPU forward and gradient kernels).

Layout of the BTK_MAC_WEIGHTS = 5644 parameters (include/btk.h): W1[64][10],
b1[64], W2[64][64], b2[64], W3[12][64], b3[12]; ReLU hidden layers; loss =
mean over the batch and the 12 outputs of (output - target)^2.  The
gradient is pinned by central finite differences in tests/test_mac_oracle.py.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 10, 64, 12
SIZES = [("W1", (N_HID, N_IN)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
         ("W3", (N_OUT, N_HID)), ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(w: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = w[o:o + k].reshape(shape)
        o += k
    return out


def forward(w: np.ndarray, X: np.ndarray):
    p = unpack(np.asarray(w, np.float64))
    a1 = X @ p["W1"].T + p["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ p["W2"].T + p["b2"]
    h2 = np.maximum(a2, 0.0)
    return h2 @ p["W3"].T + p["b3"], (a1, h1, a2, h2)


def loss_and_grad(w: np.ndarray, X: np.ndarray, Y: np.ndarray):
    p = unpack(np.asarray(w, np.float64))
    out, (a1, h1, a2, h2) = forward(w, X)
    r = out - Y
    n = X.shape[0]
    loss = float(np.mean(r * r))
    d3 = 2.0 * r / (n * N_OUT)
    d2 = (d3 @ p["W3"]) * (a2 > 0)
    d1 = (d2 @ p["W2"]) * (a1 > 0)
    g = {"W3": d3.T @ h2, "b3": d3.sum(0), "W2": d2.T @ h1, "b2": d2.sum(0), "W1": d1.T @ X, "b1": d1.sum(0)}
    return loss, np.concatenate([g[name].ravel() for name, _ in SIZES])
