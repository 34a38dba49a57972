"""Pins of the FP64 MLP oracle (oracle/mlp.py) used to check the device's
learned-warm-start kernels: finite differences for both losses' gradients,
hand-built networks for the forward pass and the start parameterisation, BF16
rounding known answers, feature special cases.  CPU only."""
import math

import numpy as np

import oracle.mlp as M


def _random_theta(rng, scale=1.0):
    p = {}
    for name, shp in M.SHAPES:
        fan_in = shp[1] if len(shp) == 2 else 128
        p[name] = rng.uniform(-1, 1, size=shp) * scale / math.sqrt(fan_in if name != "W1" else 12)
    return M.pack(p)


def test_param_count_and_order():
    assert M.N_PARAMS == 35075
    th = np.arange(M.N_PARAMS, dtype=np.float64)
    p = M.unpack(th)
    assert p["W1"][1, 0] == 12 and p["b1"][0] == 1536 and p["W4"][0, 0] == 34688
    assert np.array_equal(M.pack(p), th)


def test_bf16_round_to_nearest_even():
    eps = 2.0 ** -7  # BF16 spacing on [1, 2)
    cases = {1.0: 1.0, 1.0 + eps / 2: 1.0, 1.0 + 1.5 * eps: 1.0 + 2 * eps,
             1.0 + 0.75 * eps: 1.0 + eps, -(1.0 + eps / 2): -1.0, 3.0e38: 1.765625 * 2.0 ** 127}
    for x, want in cases.items():
        got = float(M.bf16(np.array([x]))[0])
        assert got == np.float32(want), (x, got, want)


def test_forward_single_path_closed_form():
    # unit 0 of every layer carries x0 through three SiLUs; everything else 0
    p = {name: np.zeros(shp) for name, shp in M.SHAPES}
    p["W1"][0, 0] = p["W2"][0, 0] = p["W3"][0, 0] = 1.0
    p["W4"][1, 0] = 1.0
    p["b4"][:] = [0.25, 0.0, -0.5]
    X = np.zeros((2, 12))
    X[0, 0], X[1, 0] = 1.0, -2.0
    y = M.forward(M.pack(p), X, bf16_weights=False)
    s = lambda z: z / (1.0 + math.exp(-z))
    for c in range(2):
        want = s(s(s(X[c, 0])))
        assert abs(y[c, 1] - want) < 1e-15
        assert y[c, 0] == 0.25 and y[c, 2] == -0.5


def test_forward_matches_explicit_sums():
    rng = np.random.default_rng(0)
    th = _random_theta(rng)
    X = rng.normal(size=(3, 12))
    p = M.unpack(th)
    y = M.forward(th, X, bf16_weights=False)
    for c in range(3):
        a = X[c]
        for k, (W, b) in enumerate([(p["W1"], p["b1"]), (p["W2"], p["b2"]), (p["W3"], p["b3"])]):
            a = np.array([sum(W[i, j] * a[j] for j in range(len(a))) + b[i] for i in range(128)])
            a = np.array([z / (1.0 + math.exp(-z)) for z in a])
        want = [sum(p["W4"][o, j] * a[j] for j in range(128)) + p["b4"][o] for o in range(3)]
        np.testing.assert_allclose(y[c], want, rtol=1e-12, atol=1e-14)


def _fd_check(lossfn, th, g, rng, count=60):
    offs = np.cumsum([0] + [int(np.prod(sh)) for _, sh in M.SHAPES])
    idx = list(rng.choice(M.N_PARAMS, size=count, replace=False))
    idx += [int(o) for o in offs[:-1]] + [M.N_PARAMS - 1]  # first entry of every tensor
    for i in idx:
        h = 1e-5 * max(1.0, abs(th[i]))
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (lossfn(tp) - lossfn(tm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g).max()), (i, fd, g[i])


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    th = _random_theta(rng, scale=2.0)
    X = rng.normal(size=(7, 12))
    T = np.concatenate([rng.normal(size=(7, 3)), rng.uniform(0.1, 2.0, size=(7, 1))], axis=1)
    L, g = M.loss_and_grad(th, X, T, bf16_weights=False)
    _fd_check(lambda t: M.loss_and_grad(t, X, T, False)[0], th, g, rng)
    # loss definition: weights normalised by their mean, u from the basis
    y = M.forward(th, X, bf16_weights=False)
    D1, D2 = X[:, 0:3] - X[:, 3:6], X[:, 3:6] - X[:, 6:9]
    u = y[:, :1] * D1 + y[:, 1:2] * D2 + y[:, 2:3] * X[:, 0:3]
    w = T[:, 3] / T[:, 3].mean()
    assert abs(L - np.sum(w * np.sum((u - T[:, :3]) ** 2, 1)) / 14.0) < 1e-12


def test_residual_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    th = _random_theta(rng, scale=2.0)
    n = 6
    B = rng.normal(size=(3 * n, 3 * n))
    A = B @ B.T + 3 * n * np.eye(3 * n)  # SPD like the CG operator
    steps = [(rng.normal(size=(n, 12)), np.concatenate([rng.normal(size=(n, 3)), np.ones((n, 1))], 1),
              rng.uniform(0.5, 2.0, size=n)) for _ in range(3)]
    L, g = M.residual_loss_and_grad(th, steps, A, bf16_weights=False)
    _fd_check(lambda t: M.residual_loss_and_grad(t, steps, A, False)[0], th, g, rng, count=40)
    # a network whose start is exact has zero loss; y = 0 (the plain start) has loss 1
    p0 = {name: np.zeros(sh) for name, sh in M.SHAPES}
    assert abs(M.residual_loss_and_grad(M.pack(p0), steps, A, False)[0] - 1.0) < 1e-12


def test_start_parameterisation():
    # y = (2, -1, 0) is quadratic extrapolation: v + 2 (v - v1) - (v1 - v2)
    rng = np.random.default_rng(3)
    v, v1, v2 = rng.normal(size=(3, 4, 3))
    F, s = M.features(v, v1, v2, np.ones((4, 3)), np.ones(4), np.zeros(4), 0.01, 0.1, 0.001)
    u = M.start_change(np.tile([2.0, -1.0, 0.0], (4, 1)), F)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    want = 3 * f32(v) - 3 * f32(v1) + f32(v2)
    np.testing.assert_allclose(f32(v) + s[:, None] * u, want, rtol=1e-12, atol=1e-12)


def test_features_special_cases():
    v = np.array([[3.0, 4.0, 0.0], [0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    v1 = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 2.0, 0.0]])
    v2 = np.zeros((3, 3))
    kd = np.array([[1e4, 2e4, 3e4]] * 3)
    mass = np.array([0.5, 0.5, 1e300])
    fixed = np.array([0, 0, 1])
    F, s = M.features(v, v1, v2, kd, mass, fixed, dt=0.01, alpha=0.1, beta=0.001)
    assert s[0] == 5.0 and s[1] == 0.0 and s[2] == 3.0
    np.testing.assert_allclose(F[0, :3], [0.6, 0.8, 0.0], rtol=1e-7)
    assert np.all(F[1, :9] == 0.0)
    q = (1e-4 + 1e-5) / (1.001 * 0.5)
    np.testing.assert_allclose(F[0, 9:], 0.25 * np.log1p(q * kd[0]), rtol=1e-12)
    assert np.all(F[2, 9:] == 0.0)  # fixed node: no stiffness feature
    T = M.targets(np.array([[4.0, 4.0, 0.0], [1.0, 1.0, 1.0], [0, 0, 0]]), v, s)
    np.testing.assert_allclose(T[0], [0.2, 0.0, 0.0, 25.0])
    assert np.all(T[1] == 0.0)
