"""GPU: learned CG warm start kernels against the FP64 oracle (oracle/mlp.py).

* forward pass and both loss gradients (per-sample and residual) of the
  BF16-weight MFMA network vs the oracle with the same BF16 weights (relative
  error of outputs and of each parameter tensor's gradient <= 2e-3);
* the first Adam step against its closed form (-lr * sign(g));
* recorded samples (features, targets) vs the oracle's feature map computed
  from the solver's own states;
* CG started from the network: identical to the plain start for a zero
  output layer, within the 1e-6 displacement bar of the FP64 solver oracle
  for a non-trivial network, and back to the plain start where the network
  produces NaN.
"""
import numpy as np
import pytest

from conftest import rel_disp_error

pytestmark = pytest.mark.gpu


def _theta(rng, out_scale=1.0):
    import oracle.mlp as M
    p = {}
    for name, shp in M.SHAPES:
        fan_in = 12 if name in ("W1", "b1") else 128
        p[name] = rng.uniform(-1, 1, size=shp) / np.sqrt(fan_in)
    p["W4"] *= out_scale
    p["b4"] *= out_scale
    return M.pack(p).astype(np.float32)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_matches_oracle():
    from opentissue_amd import fem
    import oracle.mlp as M
    rng = np.random.default_rng(0)
    m = fem.Model(seed=1)
    th = _theta(rng)
    m.set_params(th)
    assert np.array_equal(m.params(), th)
    X = rng.normal(size=(1000, 12)).astype(np.float32)
    y = m.forward(X)
    yo = M.forward(th.astype(np.float64), X.astype(np.float64), bf16_weights=True)
    assert _rel(y, yo) <= 2e-3, _rel(y, yo)
    assert np.abs(y - yo).max() <= 1e-2 * np.abs(yo).max()


def test_new_model_is_quadratic_extrapolation():
    from opentissue_amd import fem
    m = fem.Model(seed=3)
    y = m.forward(np.random.default_rng(0).normal(size=(64, 12)))
    assert np.all(y == np.array([2.0, -1.0, 0.0], dtype=np.float32))


@pytest.mark.parametrize("n", [128, 1000, 3000])
def test_gradient_matches_oracle(n):
    from opentissue_amd import fem
    import oracle.mlp as M
    rng = np.random.default_rng(n)
    m = fem.Model(seed=2)
    th = _theta(rng, out_scale=0.5)
    m.set_params(th)
    X = rng.normal(size=(n, 12)).astype(np.float32)
    T = np.concatenate([rng.normal(size=(n, 3)), rng.uniform(0.0, 3.0, size=(n, 1))], 1).astype(np.float32)
    loss, g = m.gradient(X, T)
    Lo, go = M.loss_and_grad(th.astype(np.float64), X.astype(np.float64), T.astype(np.float64))
    assert abs(loss - Lo) <= 1e-3 * abs(Lo), (loss, Lo)
    o = 0
    for name, shp in M.SHAPES:
        k = int(np.prod(shp))
        err = _rel(g[o:o + k], go[o:o + k])
        assert err <= 2e-3, (name, err)
        o += k


def test_first_adam_step_closed_form():
    # one sample repeated: every minibatch is that sample, and Adam's first
    # step is -lr * g / (|g| + eps) = -lr * sign(g) where |g| >> eps
    from opentissue_amd import fem
    rng = np.random.default_rng(5)
    m = fem.Model(seed=3)
    th = _theta(rng, out_scale=0.5)
    m.set_params(th)
    X = np.repeat(rng.normal(size=(1, 12)), 128, 0).astype(np.float32)
    T = np.repeat(np.array([[0.3, -0.2, 0.1, 1.5]]), 128, 0).astype(np.float32)
    loss0, g = m.gradient(X, T)
    m.reserve(128)
    m.add_samples(X, T)
    lr = 1e-3
    loss = m.train(1, batch=128, lr=lr, seed=9)
    assert abs(loss[0] - loss0) <= 1e-5 * abs(loss0)
    d = m.params().astype(np.float64) - th
    big = np.abs(g) > 1e-4
    assert big.sum() > 1000
    np.testing.assert_allclose(d[big], -lr * np.sign(g[big]), atol=1e-6)
    # the BF16 copies used by the kernels follow the FP32 masters
    y = m.forward(X[:1])
    import oracle.mlp as M
    yo = M.forward(m.params().astype(np.float64), X[:1].astype(np.float64))
    assert _rel(y, yo) <= 2e-3


def _cantilever(I, J, K):
    from opentissue_amd import mesh
    c, t = mesh.generate_blocks(I, J, K, 0.01, 0.01, 0.01)
    fx = (np.abs(c[:, 2]) < 1e-12).astype(np.uint8)
    return c, t, fx


def _rest_kdiag(c, t, fx, E=5e5, nu=0.33, rho=1000.0):
    import oracle
    o = oracle.Oracle(c, t, fx)
    o.init(E, nu, rho)
    o.simulate(0.01, True, oracle.config_params())  # from rest: every Re = I
    K = o.export_K_dense()
    return np.diag(K).reshape(-1, 3)


def test_recorded_samples_match_oracle_features():
    from opentissue_amd import fem, mesh
    import oracle.mlp as M
    c, t, fx = _cantilever(3, 6, 2)
    s = fem.Solver(c, t, fx)
    s.init(5e5, 0.33, 1000.0)
    mass = s.mass()
    f = mesh.gravity_forces(np.where(fx == 1, 0.0, mass), fx)
    f += mesh.traction_forces(c, t, 1, 1, np.array([300.0, 0.0, -500.0]))
    f[fx == 1] = 0
    s.set_external_forces(f)
    model = fem.Model(seed=0)
    free = fx == 0
    nf = int(free.sum())
    model.reserve(6 * nf)
    s.set_warmstart_model(model)
    s.set_recording(True)
    kd = _rest_kdiag(c, t, fx)
    P = fem.config_params()
    vh1 = np.zeros_like(c)
    vh2 = np.zeros_like(c)
    ref_F, ref_T = [], []
    for _ in range(6):
        _, v = s.state()
        s.simulate(0.01, True, P)
        _, v1 = s.state()
        F, sc = M.features(v, vh1, vh2, kd, mass, fx, 0.01, P.alpha, P.beta)
        ref_F.append(F[free])
        ref_T.append(M.targets(v1, v, sc)[free])
        vh2, vh1 = vh1, v
    X, T = model.samples()
    assert X.shape == (6 * nf, 12)
    RF = np.concatenate(ref_F)
    RT = np.concatenate(ref_T)
    np.testing.assert_allclose(X, RF, rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(T, RT, rtol=2e-5, atol=1e-6 * np.abs(RT[:, :3]).max())
    assert np.abs(RT[:, :3]).max() > 0 and (RT[:, 3] > 0).mean() > 0.5


def _pair_cfg(dims):
    from opentissue_amd import fem, mesh
    import oracle
    c, t, fx = _cantilever(*dims)
    s = fem.Solver(c, t, fx)
    s.init(5e5, 0.33, 1000.0)
    o = oracle.Oracle(c, t, fx)
    o.init(5e5, 0.33, 1000.0)
    m = o.mass()
    f = mesh.gravity_forces(np.where(fx == 1, 0.0, m), fx)
    f += mesh.traction_forces(c, t, 1, 1, np.array([0.0, 400.0, -800.0]))
    f[fx == 1] = 0
    s.set_external_forces(f)
    o.set_external_forces(f)
    return c, s, o


def test_zero_output_layer_is_the_plain_start():
    from opentissue_amd import fem
    import oracle.mlp as M
    c, s1, _ = _pair_cfg((3, 6, 2))
    _, s2, _ = _pair_cfg((3, 6, 2))
    model = fem.Model(seed=4)
    p = M.unpack(model.params().astype(np.float64))
    p["W4"][:] = 0.0
    p["b4"][:] = 0.0  # y = 0: the start is v^n exactly
    model.set_params(M.pack(p).astype(np.float32))
    s2.set_warmstart_model(model)
    P = fem.config_params()
    L = fem.config_params()
    L.warm_start = fem.WARM_LEARNED
    # same start and arithmetic; the per-node sums are LDS atomics whose
    # order follows wave scheduling, so two runs agree to rounding only
    # (r0.r0 cancels about 8 digits) and the iteration that crosses rel-res
    # 1e-6 moves by a few at most
    for _ in range(5):
        a = s1.simulate(0.01, True, P)
        b = s2.simulate(0.01, True, L)
        slack = max(3, int(0.05 * a.iterations))
        assert abs(a.iterations - b.iterations) <= slack and abs(a.rr0 - b.rr0) <= 1e-4 * a.rr0
    assert rel_disp_error(s2.state()[0], s1.state()[0], c) <= 1e-7


def test_learned_start_parity_with_oracle():
    # a non-trivial network changes the CG start; CG still runs to the same
    # relative residual, so the trajectory stays within the bar of the FP64
    # oracle (free running, 8 steps)
    from opentissue_amd import fem
    import oracle
    c, s, o = _pair_cfg((3, 6, 2))
    _, sp, _ = _pair_cfg((3, 6, 2))
    rng = np.random.default_rng(7)
    model = fem.Model(seed=5)
    model.set_params(_theta(rng, out_scale=0.2))
    s.set_warmstart_model(model)
    L = fem.config_params()
    L.warm_start = fem.WARM_LEARNED
    errs, moved = [], 0
    for k in range(8):
        st = s.simulate(0.01, True, L)
        sq = sp.simulate(0.01, True, fem.config_params())
        o.simulate(0.01, True, oracle.config_params())
        errs.append(rel_disp_error(s.state()[0], o.state()[0], c))
        assert st.flags == 0 and st.bodies_converged == 1
        moved += int(abs(st.rr0 - sq.rr0) > 1e-2 * sq.rr0)
    assert max(errs) <= 1e-6, errs
    assert moved >= 5  # the network did change the start


def test_learned_start_on_batched_bodies_matches_per_body_oracle():
    """One network shared by uneven bodies stepped together, each with its own
    CG: every body stays within the bar of an oracle run of that body alone
    (free running, per-body materials and tractions)."""
    from opentissue_amd import fem, mesh
    import oracle
    rng = np.random.default_rng(12)
    dims = [(2, 3, 2), (3, 2, 4), (1, 2, 2), (4, 2, 2)]
    parts, orcs, forces, fixed, E, nu = [], [], [], [], [], []
    for b, d in enumerate(dims):
        c, t = mesh.generate_blocks(*d, 0.01, 0.01, 0.01)
        c = mesh.jitter_interior(c, t, 0.01, 0.2, 40 + b)
        fx = mesh.face_nodes(c, 0, 0).astype(np.uint8)
        Eb, nub = float(np.exp(rng.uniform(np.log(1e5), np.log(1e6)))), float(rng.uniform(0.25, 0.45))
        f = mesh.traction_forces(c, t, 0, 1, rng.normal(size=3) * 1e3)
        f[fx == 1] = 0
        o = oracle.Oracle(c, t, fx)
        o.init(Eb, nub, 1000.0)
        o.set_external_forces(f)
        parts.append((c, t)); orcs.append(o); forces.append(f); fixed.append(fx)
        E.append(np.full(len(t), Eb)); nu.append(np.full(len(t), nub))
    coords, tets, no, to = mesh.pack_bodies(parts)
    s = fem.Solver(coords, tets, np.concatenate(fixed), body_node_offsets=no, body_tet_offsets=to,
                   target_tile_nodes=16)
    s.init_materials(np.concatenate(E), np.concatenate(nu), 1000.0)
    s.set_external_forces(np.concatenate(forces))
    model = fem.Model(seed=9)
    model.set_params(_theta(rng, out_scale=0.2))
    s.set_warmstart_model(model)
    L = fem.config_params()
    L.warm_start = fem.WARM_LEARNED
    worst = 0.0
    for _ in range(6):
        st = s.simulate(0.01, True, L)
        assert st.flags == 0 and st.bodies_converged == len(dims)
        x, _ = s.state()
        for b, o in enumerate(orcs):
            o.simulate(0.01, True, oracle.config_params())
            sl = slice(no[b], no[b + 1])
            worst = max(worst, rel_disp_error(x[sl], o.state()[0], parts[b][0]))
    assert worst <= 1e-6, worst


def test_nonfinite_prediction_falls_back_to_plain():
    from opentissue_amd import fem
    import oracle.mlp as M
    c, s1, _ = _pair_cfg((2, 4, 2))
    _, s2, _ = _pair_cfg((2, 4, 2))
    rng = np.random.default_rng(8)
    th = _theta(rng)
    p = M.unpack(th.astype(np.float64))
    p["b4"][1] = np.nan  # a corrupted network: every prediction is NaN
    model = fem.Model(seed=6)
    model.set_params(M.pack(p).astype(np.float32))
    s2.set_warmstart_model(model)
    P = fem.config_params()
    L = fem.config_params()
    L.warm_start = fem.WARM_LEARNED
    for _ in range(3):
        a = s1.simulate(0.01, True, P)
        b = s2.simulate(0.01, True, L)
        assert b.flags & 2
        assert abs(a.iterations - b.iterations) <= 1 and abs(a.rr0 - b.rr0) <= 1e-4 * a.rr0
    assert rel_disp_error(s2.state()[0], s1.state()[0], c) <= 1e-9


def test_learned_without_model_is_an_error():
    from opentissue_amd import fem
    c, s, _ = _pair_cfg((1, 2, 1))
    L = fem.config_params()
    L.warm_start = fem.WARM_LEARNED
    with pytest.raises(fem.OtfemError):
        s.simulate(0.01, True, L)
    with pytest.raises(fem.OtfemError):
        s.set_recording(True)


def test_residual_gradient_matches_oracle():
    from opentissue_amd import fem, mesh
    import oracle
    import oracle.mlp as M
    c, t, fx = _cantilever(3, 6, 2)
    s = fem.Solver(c, t, fx)
    s.init(5e5, 0.33, 1000.0)
    f = mesh.gravity_forces(np.where(fx == 1, 0.0, s.mass()), fx)
    f += mesh.traction_forces(c, t, 1, 1, np.array([300.0, 200.0, -500.0]))
    f[fx == 1] = 0
    s.set_external_forces(f)
    model = fem.Model(seed=7)
    rng = np.random.default_rng(11)
    th = _theta(rng, out_scale=0.3)
    model.set_params(th)
    nf = int((fx == 0).sum())
    model.reserve(6 * nf)
    s.set_warmstart_model(model)
    s.set_recording(True)
    P = fem.config_params()
    for _ in range(6):
        s.simulate(0.01, True, P)
    s.set_recording(False)
    X, T = model.samples()
    steps = [2, 5, 3]
    loss, g = model.residual_gradient(s, 0.01, steps, P)
    # the operator from the FP64 oracle: rest stiffness (every Re = I) + mass
    o = oracle.Oracle(c, t, fx)
    o.init(5e5, 0.33, 1000.0)
    o.simulate(0.01, True, oracle.config_params())
    free3 = np.repeat(fx == 0, 3)
    K = o.export_K_dense()[np.ix_(free3, free3)]
    cM, cK = 1 + P.alpha * 0.01, 0.01 ** 2 + P.beta * 0.01
    A = cK * K + cM * np.diag(np.repeat(o.mass()[fx == 0], 3))
    sd = [(X[k * nf:(k + 1) * nf].astype(np.float64), T[k * nf:(k + 1) * nf].astype(np.float64),
           np.sqrt(T[k * nf:(k + 1) * nf, 3].astype(np.float64))) for k in steps]
    Lo, go = M.residual_loss_and_grad(th.astype(np.float64), sd, A)
    assert abs(loss - Lo) <= 2e-3 * Lo, (loss, Lo)
    o_ = 0
    for name, shp in M.SHAPES:
        k = int(np.prod(shp))
        err = _rel(g[o_:o_ + k], go[o_:o_ + k])
        assert err <= 5e-3, (name, err)
        o_ += k
