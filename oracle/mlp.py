"""CPU restatement of the learned Davidson correction (TEST INFRASTRUCTURE ONLY).

The network has no counterpart in the reference: it replaces the diagonal
part of the Dvdsn correction t_i = r_i / (theta - D_i) (src/davidson.f90:482-488)
by t_i = r_i * g(x_i), g = -f1 + MLP(x).  This module restates the forward
pass, the loss and the backward pass of the HIP kernels in numpy, rounding
to BF16 at the same points as the device code:

    x  = [bf(features) | bf(E)[bucket] | 0]           (n x 32)
    h1 = bf(relu(x bf(W1)^T + b1))                     (n x 128)
    h2 = relu(h1 bf(W2)^T + b2)                        (n x 128, FP32)
    y  = h2 w3 + b3
    t  = s * (-bf(features[:, 1]) + y)      (s: the coupled residual r + c)
    L  = inv_n * sum (t - target)^2

    backward, with dy = 2 (t - target) r inv_b:
    d2 = bf(dy w3^T [h2 > 0]),  d1 = bf((d2 bf(W2)) [h1 > 0])
    dW2 = d2^T h1, dW1 = d1^T x, dE[bucket] += (d1 bf(W1))[:, 8:24]

With bf16=False every rounding is skipped and the same code is the exact
FP64 network, which the finite-difference check of the tests pins.
Accumulation here is FP64; the device accumulates in FP32 (MFMA), so the
comparison tolerance is set by FP32 accumulation and by rare BF16 rounding
ties, not by the network's arithmetic.
"""
from __future__ import annotations

import numpy as np

NF, NIN, NH, NB, NE = 8, 32, 128, 64, 16
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NIN
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH
OFF_E = OFF_B3 + 1
NPARAM = OFF_E + NB * NE


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u >> 16) & 1) + 0x7FFF
    out = (((u + r) >> 16) << 16).astype(np.uint32)
    out = np.where(special, (u & 0xFFFF0000).astype(np.uint32), out)
    return out.view(np.float32).reshape(a.shape)


def unpack(w):
    w = np.asarray(w, np.float64)
    return dict(
        W1=w[OFF_W1:OFF_B1].reshape(NH, NIN), b1=w[OFF_B1:OFF_W2], W2=w[OFF_W2:OFF_B2].reshape(NH, NH),
        b2=w[OFF_B2:OFF_W3], w3=w[OFF_W3:OFF_B3], b3=float(w[OFF_B3]), E=w[OFF_E:].reshape(NB, NE))


def features(u, r, rms, rank, c=0.0):
    """The 8 per-sample features of the device code (make_features in mlp.hip):
    u = D - theta, the clipped inverse, tanh(u/0.1), log(|u| + 1e-3), rank/3,
    r/rms, sign(u) and c/rms (c: coupling to the head-block correction)."""
    u = np.asarray(u, np.float64)
    r = np.asarray(r, np.float64)
    c = np.asarray(c, np.float64) + 0 * u
    s = np.where(u >= 0, 1.0, -1.0)
    inv = 1.0 / rms if rms > 0 else 0.0
    X = np.stack([u, 1.0 / (u + s * 0.02), np.tanh(u / 0.1), np.log(np.abs(u) + 1e-3),
                  np.asarray(rank, np.float64) / 3.0 + 0 * u, r * inv, s, c * inv], axis=1)
    return X.astype(np.float32)


def _rd(a, on):
    """BF16 rounding point (through FP32, as on the device) or the exact value."""
    return bf16(np.asarray(a, np.float32)).astype(np.float64) if on else np.asarray(a, np.float64)


def forward(w, X, bucket, r, bf=True, keep=False):
    p = unpack(w)
    X = np.asarray(X, np.float32 if bf else np.float64)
    n = X.shape[0]
    xin = np.zeros((n, NIN))
    xin[:, :NF] = _rd(X, bf)
    xin[:, NF:NF + NE] = _rd(p["E"], bf)[np.asarray(bucket)]
    W1 = _rd(p["W1"], bf)
    W2 = _rd(p["W2"], bf)
    a1 = xin @ W1.T + p["b1"]
    h1 = _rd(np.maximum(a1, 0.0), bf)
    a2 = h1 @ W2.T + p["b2"]
    h2 = np.maximum(a2, 0.0)
    y = h2 @ p["w3"] + p["b3"]
    t = np.asarray(r, np.float64) * (-xin[:, 1] + y)
    if keep:
        return t, dict(xin=xin, W1=W1, W2=W2, h1=h1, h2=h2, p=p)
    return t


def loss_grad(w, X, bucket, r, target, bf=True, inv_n=None, inv_b=None):
    """(loss, gradient as a flat NPARAM array) of the device's training step."""
    n = len(r)
    inv_n = 1.0 / n if inv_n is None else inv_n
    inv_b = 1.0 / n if inv_b is None else inv_b
    t, s = forward(w, X, bucket, r, bf=bf, keep=True)
    d = t - np.asarray(target, np.float64)
    loss = inv_n * float(np.sum(d * d))
    dy = 2.0 * d * np.asarray(r, np.float64) * inv_b
    p = s["p"]
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = dy @ s["h2"]
    g[OFF_B3] = dy.sum()
    d2 = dy[:, None] * p["w3"][None, :] * (s["h2"] > 0)
    d2 = _rd(d2, bf)
    g[OFF_B2:OFF_W3] = d2.sum(0)
    g[OFF_W2:OFF_B2] = (d2.T @ s["h1"]).reshape(-1)
    d1 = (d2 @ s["W2"]) * (s["h1"] > 0)
    d1 = _rd(d1, bf)
    g[OFF_B1:OFF_W2] = d1.sum(0)
    g[OFF_W1:OFF_B1] = (d1.T @ s["xin"]).reshape(-1)
    dx = d1 @ s["W1"]
    gE = np.zeros((NB, NE))
    np.add.at(gE, np.asarray(bucket), dx[:, NF:NF + NE])
    g[OFF_E:] = gE.reshape(-1)
    return loss, g


def init_weights(seed=0, w3_scale=1e-3):
    """He-style initialisation with the device's layout (values differ from pci_gpu_mlp_init's)."""
    rng = np.random.default_rng(seed)
    w = np.zeros(NPARAM, np.float32)
    W1 = rng.standard_normal((NH, NIN)) * np.sqrt(2.0 / (NF + NE))
    W1[:, NF + NE:] = 0.0
    w[OFF_W1:OFF_B1] = W1.reshape(-1)
    w[OFF_W2:OFF_B2] = (rng.standard_normal(NH * NH) * np.sqrt(2.0 / NH))
    w[OFF_W3:OFF_B3] = rng.standard_normal(NH) * w3_scale
    w[OFF_E:] = rng.standard_normal(NB * NE) * 0.1
    return w
