"""FP64 CPU restatement of the learned CG warm start (network, losses,
gradients, node features).  TEST INFRASTRUCTURE ONLY: the product path runs
the HIP kernels of ``opentissue_amd/csrc/mlp.hip``; this module is the
checker.

The reference has no learned component (its CG starts from ``m_velocity``,
fem_conjugate_gradients.h:48-67), so nothing here follows a reference file.
Pins: ``tests/test_mlp_oracle.py`` checks both gradients against central
finite differences of their losses and the forward pass against hand-built
networks and explicit sums.

Network (parameter order of ``otfem_model_get_params``)::

    z1 = W1 x + b1, h1 = silu(z1), ..., z3 = W3 h2 + b3, h3 = silu(z3),
    y = W4 h3 + b4,    silu(z) = z / (1 + exp(-z))

``bf16_weights=True`` evaluates the network as the device does: weights
rounded to BF16 (round to nearest even), everything else exact (the device
feeds inputs and activations to the MFMAs as BF16 hi + lo pairs, about 16
significant bits).  Gradients with respect to the FP32 master weights treat
the weight rounding as the identity, as the device's Adam does.

Start of CG from the outputs: ``u = y0 D1 + y1 D2 + y2 F0`` with
``D1 = f[0:3] - f[3:6]``, ``D2 = f[3:6] - f[6:9]``, ``F0 = f[0:3]``;
``v_start = v^n + s u``.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 12, 128, 3
SHAPES = [("W1", (N_HID, N_IN)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
          ("W3", (N_HID, N_HID)), ("b3", (N_HID,)), ("W4", (N_OUT, N_HID)), ("b4", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)  # 35075


def unpack(theta):
    theta = np.asarray(theta, dtype=np.float64)
    assert theta.shape == (N_PARAMS,)
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = theta[o:o + n].reshape(shp)
        o += n
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def bf16(a):
    """Round to BF16 (nearest even) through FP32, returned as float64."""
    f = np.asarray(a, dtype=np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, u | 0x400000, r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(theta, X, bf16_weights=True, keep=False):
    """Outputs y [n, 3] for features X [n, 12]."""
    p = unpack(theta)
    r = bf16 if bf16_weights else (lambda a: np.asarray(a, dtype=np.float64))
    W = [r(p["W1"]), r(p["W2"]), r(p["W3"]), r(p["W4"])]
    b = [p["b1"], p["b2"], p["b3"], p["b4"]]
    a = np.asarray(X, dtype=np.float64)
    acts, zs = [a], []
    for k in range(3):
        z = a @ W[k].T + b[k]
        zs.append(z)
        a = silu(z)
        acts.append(a)
    y = a @ W[3].T + b[3]
    if keep:
        return y, dict(acts=acts, zs=zs, W=W)
    return y


def basis(X):
    """(D1, D2, F0) of each feature row, each [n, 3]."""
    X = np.asarray(X, dtype=np.float64)
    return X[:, 0:3] - X[:, 3:6], X[:, 3:6] - X[:, 6:9], X[:, 0:3]


def start_change(y, X):
    D1, D2, F0 = basis(X)
    return y[:, 0:1] * D1 + y[:, 1:2] * D2 + y[:, 2:3] * F0


def backward(cache, gy):
    """dL/dtheta from dL/dy [n, 3] and a forward cache."""
    acts, zs, W = cache["acts"], cache["zs"], cache["W"]
    grads = {"W4": gy.T @ acts[3], "b4": gy.sum(0)}
    delta = (gy @ W[3]) * dsilu(zs[2])            # dL/dz3
    for k in (2, 1, 0):
        grads[f"W{k + 1}"] = delta.T @ acts[k]
        grads[f"b{k + 1}"] = delta.sum(0)
        if k > 0:
            delta = (delta @ W[k]) * dsilu(zs[k - 1])
    return pack(grads)


def _dy_from_du(du, X):
    D1, D2, F0 = basis(X)
    return np.stack([(du * D1).sum(1), (du * D2).sum(1), (du * F0).sum(1)], axis=1)


def loss_and_grad(theta, X, T, bf16_weights=True):
    """Per-sample loss L = sum_c (w_c / mean w) |u_c - u*_c|^2 / (2 n) and
    dL/dtheta.  T is [n, 4]: target u* (3) and loss weight w."""
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    n = X.shape[0]
    w = T[:, 3] / T[:, 3].mean()
    y, cache = forward(theta, X, bf16_weights, keep=True)
    d = start_change(y, X) - T[:, :3]
    L = float(np.sum(w * np.sum(d * d, axis=1)) / (2.0 * n))
    gy = _dy_from_du((w[:, None] * d) / n, X)
    return L, backward(cache, gy)


def residual_loss_and_grad(theta, steps, A, bf16_weights=True):
    """Residual loss over recorded steps and dL/dtheta.

    steps: list of (X [n, 12], T [n, 4], s [n]) of one mesh's free nodes;
    A: [3n, 3n] operator (dense or scipy sparse).  With e = s (u - u*) and
    e0 = -s u* (velocity units), L = sum_k |A e_k|^2 / sum_k |A e0_k|^2."""
    parts, n0 = [], 0.0
    for X, T, s in steps:
        X = np.asarray(X, dtype=np.float64)
        s = np.asarray(s, dtype=np.float64)
        y, cache = forward(theta, X, bf16_weights, keep=True)
        ut = np.asarray(T, dtype=np.float64)[:, :3]
        e = (s[:, None] * (start_change(y, X) - ut)).ravel()
        q0 = A @ (-s[:, None] * ut).ravel()
        n0 += float(q0 @ q0)
        parts.append((X, s, cache, A @ e))
    if n0 <= 0.0:
        return 0.0, np.zeros(N_PARAMS)
    L, g = 0.0, np.zeros(N_PARAMS)
    for X, s, cache, q in parts:
        L += float(q @ q) / n0
        du = s[:, None] * ((2.0 / n0) * (A.T @ q)).reshape(-1, 3)
        g += backward(cache, _dy_from_du(du, X))
    return L, g


def features(v, v1, v2, kdiag, mass, fixed, dt, alpha, beta):
    """Node features and scale s of the learned warm start.

    v, v1, v2: velocities of the last three steps [N, 3] (the device keeps
    v1, v2 in FP32); kdiag: diagonal of the rest stiffness [N, 3]; mass [N].
    Returns (F [N, 12], s [N])."""
    cM = 1.0 + alpha * dt
    cK = dt * dt + beta * dt
    v, v1, v2 = (np.asarray(a, dtype=np.float32).astype(np.float64) for a in (v, v1, v2))
    s = (np.linalg.norm(v, axis=1) + np.linalg.norm(v1, axis=1) + np.linalg.norm(v2, axis=1))
    inv = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 0.0)
    F = np.zeros((len(v), 12))
    F[:, 0:3] = v * inv[:, None]
    F[:, 3:6] = v1 * inv[:, None]
    F[:, 6:9] = v2 * inv[:, None]
    free = np.asarray(fixed) == 0
    q = np.where(free, cK / (cM * np.where(free, mass, 1.0)), 0.0)
    F[:, 9:12] = 0.25 * np.log1p(q[:, None] * kdiag)
    F[~free, 9:12] = 0.0
    return F, s


def targets(v_next, v, s):
    """(v^{n+1} - v^n) / s and weight s^2; zero where s == 0."""
    v = np.asarray(v, dtype=np.float32).astype(np.float64)
    T = np.zeros((len(s), 4))
    ok = s > 0
    T[ok, :3] = (np.asarray(v_next)[ok] - v[ok]) / s[ok, None]
    T[ok, 3] = s[ok] ** 2
    return T
