"""Checkpoint / resume on the GPU: a solver's state file restores every array
bit for bit into a new solver of the same mesh, the resumed trajectory stays
with the original (per-node sums are LDS atomics, so two runs agree to
rounding), a file of another mesh is refused, and a model file restores the
parameters and Adam state exactly."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(seed=3):
    from opentissue_amd import fem, mesh
    c, t = mesh.generate_blocks(5, 9, 3, 0.01, 0.01, 0.01)
    fx = (np.abs(c[:, 2]) < 1e-12).astype(np.uint8)
    rng = np.random.default_rng(seed)
    E = np.exp(rng.uniform(np.log(2e5), np.log(8e5), len(t)))
    s = fem.Solver(c, t, fx)
    s.init_materials(E, 0.3, 1000.0, c_yield=0.0, c_creep=0.5, c_max=0.1)
    f = mesh.gravity_forces(np.where(fx == 1, 0.0, mesh.lumped_mass(c, t, 1000.0)), fx)
    f += mesh.traction_forces(c, t, 1, 1, np.array([0.0, 2e3, -4e3]))
    f[fx == 1] = 0
    s.set_external_forces(f)
    return s, c, t, fx


def test_solver_state_round_trip_and_resume(tmp_path):
    from opentissue_amd import fem
    a, c, t, fx = _setup()
    a.set_warmstart_model(fem.Model(seed=5))
    P = fem.config_params()
    for _ in range(4):
        a.simulate(0.01, True, P)
    path = str(tmp_path / "state.bin")
    a.save_state(path)
    xa, va = a.state()
    pa, ita = a.plastic(), a.iteration_totals()
    assert np.abs(pa).max() > 0  # plasticity is active

    b, _, _, _ = _setup()
    b.set_warmstart_model(fem.Model(seed=5))
    b.load_state(path)
    xb, vb = b.state()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    assert np.array_equal(pa, b.plastic())
    assert np.array_equal(ita, b.iteration_totals())

    L = fem.config_params()
    L.warm_start = fem.WARM_LEARNED
    for _ in range(3):
        a.simulate(0.01, True, L)
        b.simulate(0.01, True, L)
    xa, _ = a.state()
    xb, _ = b.state()
    disp = np.abs(xa - c).max()
    assert np.abs(xa - xb).max() <= 1e-8 * disp


def test_state_of_another_mesh_is_refused(tmp_path):
    from opentissue_amd import fem, mesh
    a, _, _, _ = _setup()
    path = str(tmp_path / "state.bin")
    a.save_state(path)
    c, t = mesh.generate_blocks(5, 9, 4, 0.01, 0.01, 0.01)
    other = fem.Solver(c, t, np.zeros(len(c), np.uint8))
    other.init(5e5, 0.3, 1000.0)
    with pytest.raises(fem.OtfemError):
        other.load_state(path)
    with open(path, "rb") as f:
        head = f.read(200)
    bad = str(tmp_path / "short.bin")
    with open(bad, "wb") as f:
        f.write(head)
    b, _, _, _ = _setup()
    with pytest.raises(fem.OtfemError):
        b.load_state(bad)


def test_model_file_round_trip(tmp_path):
    from opentissue_amd import fem
    rng = np.random.default_rng(0)
    m1 = fem.Model(seed=1)
    m1.reserve(4096)
    X = rng.normal(size=(4096, fem.MLP_INPUTS)).astype(np.float32)
    T = np.concatenate([rng.normal(size=(4096, 3)) * 0.1, np.ones((4096, 1))], 1).astype(np.float32)
    m1.add_samples(X, T)
    m1.train(5, batch=1024, lr=1e-3, seed=2)
    path = str(tmp_path / "model.bin")
    m1.save(path)
    m2 = fem.Model(seed=9)
    m2.reserve(4096)
    m2.add_samples(X, T)
    m2.load(path)
    assert np.array_equal(m1.params(), m2.params())
    # same Adam state and step count: the next updates agree
    m1.train(3, batch=1024, lr=1e-3, seed=7)
    m2.train(3, batch=1024, lr=1e-3, seed=7)
    assert np.allclose(m1.params(), m2.params(), rtol=0, atol=1e-6)
    with pytest.raises(fem.OtfemError):
        m2.load(str(tmp_path / "missing.bin"))


def test_resume_with_the_subspace_history(tmp_path):
    """The subspace start's history (velocity differences, v^n-1, ring
    position) travels in the state file: in the deterministic mode a resumed
    run gives the same bits as the uninterrupted one, step for step."""
    from opentissue_amd import fem
    a, c, t, fx = _setup()
    a.set_deterministic(True)
    P = fem.config_params(subspace=6)
    for _ in range(9):  # ring of 6 slots wrapped around
        a.simulate(0.01, True, P)
    path = str(tmp_path / "state_sub.bin")
    a.save_state(path)
    b, _, _, _ = _setup()
    b.set_deterministic(True)
    b.load_state(path)
    for _ in range(4):
        sa = a.simulate(0.01, True, P)
        sb = b.simulate(0.01, True, P)
        assert sa.iterations == sb.iterations
    xa, va = a.state()
    xb, vb = b.state()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    assert np.array_equal(a.subspace_coefficients(), b.subspace_coefficients())


def test_rejected_state_file_leaves_the_solver_unchanged(tmp_path):
    """A state file whose rest shape inverts a tetrahedron is refused, and the
    solver keeps its rest shape, fixed flags, materials and masses: its next
    step is the one it would have taken."""
    import struct
    from opentissue_amd import fem
    a, c, t, fx = _setup()
    a.set_deterministic(True)  # so that "the step it would have taken" is one set of bits
    P = fem.config_params()
    a.simulate(0.01, True, P)
    path = str(tmp_path / "state.bin")
    a.save_state(path)
    # flip the rest position of one corner of tet 0 through its opposite face
    N = len(c)
    with open(path, "r+b") as f:
        off = 8 + 4 + 4 + 24 + 8  # magic, version, flags, dims, hash -> x0
        node = int(t[0, 0])
        f.seek(off + 24 * node)
        x0 = np.array(struct.unpack("3d", f.read(24)))
        cen = c[t[0, 1:]].mean(0)
        f.seek(off + 24 * node)
        f.write(struct.pack("3d", *(2 * cen - x0)))
    m0 = a.mass()
    ref, _, _, _ = _setup()
    ref.set_deterministic(True)
    ref.simulate(0.01, True, P)
    with pytest.raises(fem.OtfemError):
        a.load_state(path)
    assert np.array_equal(a.mass(), m0)
    a.simulate(0.01, True, P)
    ref.simulate(0.01, True, P)
    xa, va = a.state()
    xr, vr = ref.state()
    assert np.array_equal(xa, xr) and np.array_equal(va, vr)
