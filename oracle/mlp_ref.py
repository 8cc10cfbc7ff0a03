"""CPU reference of the learned-boundary network -- test infrastructure only.

The network has no counterpart in the reference project (which ends every
face with a CPML, src/layout.rs:471-513); this module restates the model the
GPU kernels implement (wavefront_amd/csrc/wf_learned.hip) so the tests can
check them:

  forward_f64     the network as a real function: f64 arithmetic, the
                  weights as the GPU uses them (rounded to bf16), no rounding
                  of activations;
  forward_bf16    the same with the GPU's rounding points emulated (bf16
                  inputs and hidden activations, f32 accumulation);

The hidden activation is hardtanh, clip(z, -1, 1), with derivative 1 inside
(-1, 1) and 0 where it clips.
  grad_f64        the exact gradient of the mean squared error of forward_f64
                  with respect to the FP32 master parameters (straight
                  through the bf16 rounding of the weights), checked against
                  central finite differences in the tests;
  *_one_way_f64   the same for the one-way output head;
  *_delta_f64     the same for the delta head.

Parameter layout: W1[16][64] b1[64] W2[64][64] b2[64] W3[64] b3 wlin[16]
(W[in][out]).
"""
from __future__ import annotations

import numpy as np

IN, HID = 16, 64
SIZES = (("W1", IN * HID), ("b1", HID), ("W2", HID * HID), ("b2", HID), ("W3", HID), ("b3", 1),
         ("wlin", IN))
P = sum(n for _, n in SIZES)


def split(p):
    out, o = {}, 0
    for name, n in SIZES:
        out[name] = np.asarray(p[o:o + n])
        o += n
    out["W1"] = out["W1"].reshape(IN, HID)
    out["W2"] = out["W2"].reshape(HID, HID)
    return out


def join(d):
    return np.concatenate([np.asarray(d[name], np.float64).reshape(-1) for name, _ in SIZES])


def bf16(x):
    """Round-to-nearest-even to bfloat16, returned as float32 values."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def _gpu_weights(p):
    w = split(np.asarray(p, np.float32))
    return {"W1": bf16(w["W1"]).astype(np.float64), "W2": bf16(w["W2"]).astype(np.float64),
            "W3": bf16(w["W3"]).astype(np.float64), "b1": w["b1"].astype(np.float64),
            "b2": w["b2"].astype(np.float64), "b3": float(w["b3"][0]),
            "wlin": w["wlin"].astype(np.float64)}


def act(z):
    return np.clip(z, -1.0, 1.0)


def dact(a):
    """Derivative of the activation from its value."""
    return (np.abs(a) < 1.0).astype(np.float64)


def forward_f64(p, u):
    w = _gpu_weights(p)
    u = np.asarray(u, np.float64)
    a1 = act(u @ w["W1"] + w["b1"])
    a2 = act(a1 @ w["W2"] + w["b2"])
    return a2 @ w["W3"] + w["b3"] + u @ w["wlin"], (a1, a2)


def forward_bf16(p, u):
    """The GPU's rounding points: bf16 inputs and activations into each
    matrix product, f32 accumulation starting from the bias (products of
    bf16 values are exact in f32), f32 activation, f32 linear skip."""
    w = _gpu_weights(p)
    u32 = np.asarray(u, np.float32)
    ub = bf16(u32).astype(np.float64)
    b1 = w["b1"].astype(np.float32).astype(np.float64)
    b2 = w["b2"].astype(np.float32).astype(np.float64)
    z1 = (ub @ w["W1"] + b1).astype(np.float32)
    a1 = np.clip(z1, np.float32(-1), np.float32(1))
    z2 = (bf16(a1).astype(np.float64) @ w["W2"] + b2).astype(np.float32)
    a2 = np.clip(z2, np.float32(-1), np.float32(1))
    z3 = (bf16(a2).astype(np.float64) @ w["W3"]).astype(np.float32)
    wl = w["wlin"].astype(np.float32)
    lin = (u32 * wl).astype(np.float32)
    parts = [lin[:, 4 * q:4 * q + 4].sum(axis=1, dtype=np.float32) for q in range(4)]
    s = (parts[0] + parts[1]) + (parts[2] + parts[3])
    return (z3 + np.float32(w["b3"])) + s


def loss_f64(p, u, t):
    y, _ = forward_f64(p, u)
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))


def one_way_sign(u):
    """Sign normalizing the one-way head's input: sign of u0, or of u4 when
    u0 = 0."""
    u = np.asarray(u)
    return np.where((u[:, 0] > 0) | ((u[:, 0] == 0) & (u[:, 4] >= 0)), 1.0, -1.0)


def predict_one_way_f64(p, u):
    """One-way head: u14 + tanh(y(sigma u)) (u0 - u12)."""
    u = np.asarray(u, np.float64)
    y, _ = forward_f64(p, u * one_way_sign(u)[:, None])
    return u[:, 14] + np.tanh(y) * (u[:, 0] - u[:, 12])


def loss_one_way_f64(p, u, t):
    return float(np.mean((predict_one_way_f64(p, u) - np.asarray(t, np.float64)) ** 2))


def grad_one_way_f64(p, u, t):
    """d loss / d p for the one-way head: the direct-head gradient of the
    network on sigma*u with the per-sample error scaled by
    (1 - tanh^2 y) (u0 - u12)."""
    u = np.asarray(u, np.float64)
    ui = u * one_way_sign(u)[:, None]
    y, _ = forward_f64(p, ui)
    th = np.tanh(y)
    pred = u[:, 14] + th * (u[:, 0] - u[:, 12])
    dy = (pred - np.asarray(t, np.float64)) * (1.0 - th * th) * (u[:, 0] - u[:, 12])
    # grad_f64 differentiates mean((y - t')^2); choose t' = y - dy so its error is dy
    return grad_f64(p, ui, y - dy)


def _weights(p, rounded):
    if rounded:
        return _gpu_weights(p)
    w = split(np.asarray(p, np.float64))
    w["b3"] = float(w["b3"][0])
    return w


def _mlp_f64(p, v, rounded=True):
    """The MLP part alone (no linear skip): a2 @ W3 + b3."""
    w = _weights(p, rounded)
    a1 = act(v @ w["W1"] + w["b1"])
    a2 = act(a1 @ w["W2"] + w["b2"])
    return a2 @ w["W3"] + w["b3"], (a1, a2)


def predict_delta_f64(p, u, rounded=True):
    """Delta head: u @ wlin + m(sigma u) (u0 - u14), m = the MLP part.
    rounded=False: unrounded weights (for finite differences)."""
    u = np.asarray(u, np.float64)
    w = _weights(p, rounded)
    m, _ = _mlp_f64(p, u * one_way_sign(u)[:, None], rounded)
    return u @ w["wlin"] + m * (u[:, 0] - u[:, 14])


def loss_delta_f64(p, u, t, rounded=True):
    return float(np.mean((predict_delta_f64(p, u, rounded) - np.asarray(t, np.float64)) ** 2))


def grad_delta_f64(p, u, t, rounded=True):
    """d loss / d p for the delta head: the skip's gradient from the error
    itself, the MLP's from the error times (u0 - u14)."""
    u = np.asarray(u, np.float64)
    n = len(u)
    w = _weights(p, rounded)
    v = u * one_way_sign(u)[:, None]
    m, (a1, a2) = _mlp_f64(p, v, rounded)
    d = u[:, 0] - u[:, 14]
    e = 2.0 * (u @ w["wlin"] + m * d - np.asarray(t, np.float64)) / n
    gm = e * d
    out = {"wlin": u.T @ e, "b3": np.array([gm.sum()]), "W3": a2.T @ gm}
    dz2 = np.outer(gm, w["W3"]) * dact(a2)
    out["W2"] = a1.T @ dz2
    out["b2"] = dz2.sum(0)
    dz1 = (dz2 @ w["W2"].T) * dact(a1)
    out["W1"] = v.T @ dz1
    out["b1"] = dz1.sum(0)
    return join(out)


def grad_f64(p, u, t):
    """d loss / d p for loss = mean((y - t)^2)."""
    w = _gpu_weights(p)
    u = np.asarray(u, np.float64)
    n = len(u)
    y, (a1, a2) = forward_f64(p, u)
    g = 2.0 * (y - np.asarray(t, np.float64)) / n
    d = {"b3": np.array([g.sum()]), "W3": a2.T @ g, "wlin": u.T @ g}
    dz2 = np.outer(g, w["W3"]) * dact(a2)
    d["W2"] = a1.T @ dz2
    d["b2"] = dz2.sum(0)
    dz1 = (dz2 @ w["W2"].T) * dact(a1)
    d["W1"] = u.T @ dz1
    d["b1"] = dz1.sum(0)
    return join(d)


def grad_fd(p, u, t, idx, h=1e-4):
    """Central finite differences of loss_f64 in f64 for parameter indices
    `idx`, with the bf16 rounding of the weights switched off (so the
    derivative exists): used only to pin grad_f64's formulas."""
    out = []
    for i in idx:
        q1 = np.array(p, np.float64)
        q2 = np.array(p, np.float64)
        q1[i] += h
        q2[i] -= h
        out.append((_loss_exact(q1, u, t) - _loss_exact(q2, u, t)) / (2 * h))
    return np.array(out)


def grad_exact(p, u, t):
    """grad_f64's formulas with unrounded f64 weights (for the FD check)."""
    w = split(np.asarray(p, np.float64))
    u = np.asarray(u, np.float64)
    n = len(u)
    a1 = act(u @ w["W1"] + w["b1"])
    a2 = act(a1 @ w["W2"] + w["b2"])
    y = a2 @ w["W3"] + w["b3"][0] + u @ w["wlin"]
    g = 2.0 * (y - np.asarray(t, np.float64)) / n
    d = {"b3": np.array([g.sum()]), "W3": a2.T @ g, "wlin": u.T @ g}
    dz2 = np.outer(g, w["W3"]) * dact(a2)
    d["W2"] = a1.T @ dz2
    d["b2"] = dz2.sum(0)
    dz1 = (dz2 @ w["W2"].T) * dact(a1)
    d["W1"] = u.T @ dz1
    d["b1"] = dz1.sum(0)
    return join(d)


def _loss_exact(p, u, t):
    w = split(np.asarray(p, np.float64))
    u = np.asarray(u, np.float64)
    a1 = act(u @ w["W1"] + w["b1"])
    a2 = act(a1 @ w["W2"] + w["b2"])
    y = a2 @ w["W3"] + w["b3"][0] + u @ w["wlin"]
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))
