"""CPU reference of the learned-boundary network -- test infrastructure only.

The network has no counterpart in the reference project (which ends every
face with a CPML, src/layout.rs:471-513); this module restates the model the
GPU kernels implement (wavefront_amd/csrc/wf_learned.hip) so the tests can
check them:

  forward_f64     the network as a real function: f64 arithmetic, the
                  weights as the GPU uses them (rounded to bf16), no rounding
                  of activations;
  forward_bf16    the same with the GPU's rounding points emulated (inputs
                  and hidden activations as bf16 pairs, f32 accumulation);

The hidden activation is hardtanh, clip(z, -1, 1), with derivative 1 inside
(-1, 1) and 0 where it clips.
  grad_f64        the exact gradient of the mean squared error of forward_f64
                  with respect to the FP32 master parameters (straight
                  through the bf16 rounding of the weights), checked against
                  central finite differences in the tests;
  *_one_way_f64   the same for the one-way output head;
  *_delta_f64     the same for the delta head;
  *_higdon_f64    the same for the Higdon head (higdon_terms: its product
                  of one-way factors expanded independently);
  predict_diff_f64  the difference head.

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


def bf16_pair(x):
    """x as the sum of two bfloat16 values, hi + lo (hi = x rounded, lo =
    the remainder rounded), returned as float64."""
    x = np.asarray(x, np.float32)
    hi = bf16(x)
    lo = bf16(x - hi)
    return hi.astype(np.float64) + lo.astype(np.float64)


def forward_bf16(p, u):
    """The GPU's rounding points: inputs and activations into each matrix
    product as bf16 pairs (hi + lo) against the bf16 weights, f32
    accumulation starting from the bias (products of bf16 values are exact
    in f32), f32 activation, f32 linear skip."""
    w = _gpu_weights(p)
    u32 = np.asarray(u, np.float32)
    ub = bf16_pair(u32)
    b1 = w["b1"].astype(np.float32).astype(np.float64)
    b2 = w["b2"].astype(np.float32).astype(np.float64)
    z1 = (ub @ w["W1"] + b1).astype(np.float32)
    a1 = np.clip(z1, np.float32(-1), np.float32(1))
    z2 = (bf16_pair(a1) @ w["W2"] + b2).astype(np.float32)
    a2 = np.clip(z2, np.float32(-1), np.float32(1))
    z3 = (bf16_pair(a2) @ w["W3"]).astype(np.float32)
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


def higdon_terms(kf, gf, ga, u):
    """(A, B) of the Higdon head, ghost = A + ka B, from the product of the
    two one-way factors as polynomials in X (one cell inward) and T (one
    step back): c(X, T) = 1 + k g T - X (g T + k) each; the product's
    coefficients p[x][t] give ghost = -sum p[x][t] W_x^{n-t} over
    (x, t) != (0, 0), with W_0 the ghost, W_1 = P0, W_2 = P1 (inputs: 0 P0,
    1 P1, 10, 11 P1(n-1), P1(n-2), 12, 13 g(n-1), g(n-2), 14, 15 P0(n-1),
    P0(n-2)).  A is the part at ka = 0, B the coefficient of ka."""
    u = np.asarray(u, np.float64)
    where = {(0, 1): 12, (0, 2): 13, (1, 0): 0, (1, 1): 14, (1, 2): 15, (2, 0): 1, (2, 1): 10,
             (2, 2): 11}

    def factor(k, g):
        return np.array([[1.0, k * g], [-k, -g]])

    def ghost(ka):
        a, b = factor(kf, gf), factor(ka, ga)
        p = np.zeros((3, 3))
        for i in range(2):
            for j in range(2):
                p[i:i + 2, j:j + 2] += a[i, j] * b
        return -sum(p[x, t] * u[:, i] for (x, t), i in where.items()) / p[0, 0]

    A = ghost(0.0)
    return A, ghost(1.0) - A


def _higdon_fixed(p):
    wl = split(np.asarray(p, np.float32))["wlin"].astype(np.float64)
    return wl[0], wl[1], wl[2]


def predict_higdon_f64(p, u, rounded=True):
    """Higdon head: A + tanh(y) B, y = the MLP part on the sign-normalized
    inputs; kf, gf, ga = wlin[0..2]."""
    u = np.asarray(u, np.float64)
    kf, gf, ga = _higdon_fixed(p)
    y, _ = _mlp_f64(p, u * one_way_sign(u)[:, None], rounded)
    A, B = higdon_terms(kf, gf, ga, u)
    return A + np.tanh(y) * B


def loss_higdon_f64(p, u, t, rounded=True):
    return float(np.mean((predict_higdon_f64(p, u, rounded) - np.asarray(t, np.float64)) ** 2))


def grad_higdon_f64(p, u, t, rounded=True):
    """d loss / d p for the Higdon head: the network's parameters from the
    error times (1 - ka^2) B; the skip weights (the fixed factor) get
    none."""
    u = np.asarray(u, np.float64)
    n = len(u)
    kf, gf, ga = _higdon_fixed(p)
    w = _weights(p, rounded)
    v = u * one_way_sign(u)[:, None]
    y, (a1, a2) = _mlp_f64(p, v, rounded)
    A, B = higdon_terms(kf, gf, ga, u)
    ka = np.tanh(y)
    e = 2.0 * (A + ka * B - np.asarray(t, np.float64)) / n
    gm = e * (1.0 - ka * ka) * B
    out = {"wlin": np.zeros(IN), "b3": np.array([gm.sum()]), "W3": a2.T @ gm}
    dz2 = np.outer(gm, w["W3"]) * dact(a2)
    out["W2"] = a1.T @ dz2
    out["b2"] = dz2.sum(0)
    dz1 = (dz2 @ w["W2"].T) * dact(a1)
    out["W1"] = v.T @ dz1
    out["b1"] = dz1.sum(0)
    return join(out)


def diff_features(u):
    """The difference head's inputs: P and Q depth steps, inputs 8, 9 as
    they are, the P1 and ghost / P0 history as differences (3 and 7 zero)."""
    u = np.asarray(u, np.float64)
    d = np.zeros_like(u)
    for b in (0, 4):
        d[:, b:b + 3] = u[:, b:b + 3] - u[:, b + 1:b + 4]
    d[:, 8:10] = u[:, 8:10]
    d[:, 10] = u[:, 10] - u[:, 1]
    d[:, 11] = u[:, 11] - u[:, 10]
    d[:, 12] = u[:, 12] - u[:, 0]
    d[:, 13] = u[:, 13] - u[:, 12]
    d[:, 14] = u[:, 14] - u[:, 0]
    d[:, 15] = u[:, 15] - u[:, 14]
    return d


def predict_diff_f64(p, u):
    """Difference head: u0 + (network + skip) on the differences."""
    u = np.asarray(u, np.float64)
    y, _ = forward_f64(p, diff_features(u))
    return u[:, 0] + y
