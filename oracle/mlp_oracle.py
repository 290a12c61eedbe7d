"""CPU ORACLE of the surrogate network (TEST INFRASTRUCTURE ONLY).

Forward and backward of y = W4 relu(W3 relu(W2 relu(W1 x + b1) + b2) + b3) + b4
with the rounding points of the BF16-MFMA kernels (ultraballpit_amd/csrc/
ubp_mlp.hip): inputs normalised in FP32 then rounded to BF16, weights rounded
to BF16, products accumulated here in FP64 (the device accumulates in FP32),
bias added in FP32, activations and back-propagated errors rounded to BF16
(round to nearest even).  With ``rounding=False`` everything stays FP64 and
unrounded, which is the mode the finite-difference gradient checks pin.

The network has no counterpart in the reference; its oracle is pinned only by
its own finite-difference gradient checks and by BF16 rounding checked
against an independent implementation (torch's float -> bfloat16 cast).
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 32, 128, 6
PARAMS = (("w1", (IN, HID)), ("b1", (HID,)), ("w2", (HID, HID)), ("b2", (HID,)),
          ("w3", (HID, HID)), ("b3", (HID,)), ("w4", (HID, OUT)), ("b4", (OUT,)))
STATS = (("in_mean", (IN,)), ("in_inv_std", (IN,)), ("out_mean", (OUT,)), ("out_std", (OUT,)))
NPARAM = sum(int(np.prod(s)) for _, s in PARAMS)


def bf16(a) -> np.ndarray:
    """Round FP32 values to BF16 (nearest even), returned as FP32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def init_model(seed: int = 0, scale: float = 1.0) -> dict:
    """He-style random weights, zero biases, identity normalisation."""
    rng = np.random.default_rng(seed)
    m = {}
    for name, shape in PARAMS:
        if name.startswith("w"):
            m[name] = (rng.normal(0.0, scale * np.sqrt(2.0 / shape[0]), shape)).astype(np.float32)
        else:
            m[name] = np.zeros(shape, np.float32)
    m["in_mean"] = np.zeros(IN, np.float32)
    m["in_inv_std"] = np.ones(IN, np.float32)
    m["out_mean"] = np.zeros(OUT, np.float32)
    m["out_std"] = np.ones(OUT, np.float32)
    return m


def flat(m: dict) -> np.ndarray:
    return np.concatenate([np.asarray(m[k], np.float64).ravel() for k, _ in PARAMS])


def unflat(v: np.ndarray, like: dict) -> dict:
    out = dict(like)
    o = 0
    for k, s in PARAMS:
        n = int(np.prod(s))
        out[k] = np.asarray(v[o:o + n]).reshape(s)
        o += n
    return out


def _norm_input(m, x, rounding):
    if rounding:
        xn = (np.asarray(x, np.float32) - m["in_mean"].astype(np.float32)) * m["in_inv_std"].astype(np.float32)
        return bf16(xn).astype(np.float64)
    return (np.asarray(x, np.float64) - m["in_mean"]) * m["in_inv_std"]


def forward(m: dict, x, rounding: bool = True, keep: bool = False):
    """Network output y (rows x OUT).  keep=True also returns the activations."""
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    W = [np.asarray(r(m[f"w{k}"]), np.float64) for k in (1, 2, 3, 4)]
    B = [np.asarray(m[f"b{k}"], np.float32 if rounding else np.float64) for k in (1, 2, 3, 4)]
    h = _norm_input(m, x, rounding)
    acts = [h]
    for k in range(3):
        z = h @ W[k]
        if rounding:
            z = z.astype(np.float32) + B[k]
            h = bf16(np.maximum(z, np.float32(0))).astype(np.float64)
        else:
            z = z + B[k]
            h = np.maximum(z, 0.0)
        acts.append(h)
    y = h @ W[3]
    y = (y.astype(np.float32) + B[3]) if rounding else (y + B[3])
    return (y, acts) if keep else y


def loss_grad(m: dict, x, t, rounding: bool = True):
    """L = sum (y - t)^2 / (OUT rows) and dL/dparams (dict, FP64)."""
    rows = len(x)
    y, acts = forward(m, x, rounding, keep=True)
    t = np.asarray(t, np.float32 if rounding else np.float64)
    e = y - t
    L = float(np.sum(np.asarray(e, np.float64) ** 2) / (float(OUT) * rows))
    if rounding:
        scale = np.float32(2.0) / (np.float32(OUT) * np.float32(rows))
        dz = bf16(e.astype(np.float32) * scale).astype(np.float64)
        W = [bf16(m[f"w{k}"]).astype(np.float64) for k in (1, 2, 3, 4)]
    else:
        dz = e * (2.0 / (float(OUT) * rows))
        W = [np.asarray(m[f"w{k}"], np.float64) for k in (1, 2, 3, 4)]
    g = {}
    for k in (4, 3, 2, 1):
        h = acts[k - 1]
        g[f"w{k}"] = h.T @ dz
        g[f"b{k}"] = dz.sum(axis=0)
        if k > 1:
            dh = dz @ W[k - 1].T
            mask = acts[k - 1] > 0
            dz = (bf16(dh.astype(np.float32)).astype(np.float64) if rounding else dh) * mask
    return L, g


def predict_accel(m: dict, feats, gmag):
    """Engine outputs (a_x, a_v): |g| (out_mean + out_std y) split into the
    position and velocity updates."""
    y = forward(m, feats)
    a = np.float32(gmag) * (m["out_mean"] + m["out_std"] * y)
    return a[:, :3], a[:, 3:]


def adam_step(m: dict, g: dict, state: dict, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step in FP32 (the update rule of the device optimiser)."""
    t = state.get("t", 0) + 1
    state["t"] = t
    c1 = np.float32(1.0 - beta1 ** t)
    c2 = np.float32(1.0 - beta2 ** t)
    out = dict(m)
    for k, _ in PARAMS:
        gk = np.asarray(g[k], np.float32)
        a = np.float32(beta1) * state.setdefault("m_" + k, np.zeros_like(gk)) + np.float32(1.0 - beta1) * gk
        b = np.float32(beta2) * state.setdefault("v_" + k, np.zeros_like(gk)) + np.float32(1.0 - beta2) * gk * gk
        state["m_" + k], state["v_" + k] = a, b
        out[k] = (np.asarray(m[k], np.float32) - np.float32(lr) * (a / c1) / (np.sqrt(b / c2) + np.float32(eps))
                  ).astype(np.float32)
    return out
