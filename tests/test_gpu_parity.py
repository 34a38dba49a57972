"""GPU parity: the HIP path through the C ABI against the FP64 CPU oracle.

Error metric (per step, from a shared state): after setting the GPU solver to
the oracle's (x, v) and taking one step on both,
    e = max_i |x_gpu - x_oracle| / max_i |x_oracle - x0|.
The bar is 1e-6, the per-step tolerance this floating-point path is held to.
"""
import numpy as np
import pytest

from conftest import rel_disp_error

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _lattice(I, J, K, h=0.01):
    from opentissue_amd import mesh
    c, t = mesh.generate_blocks(I, J, K, h, h, h)
    fixed = (np.abs(c[:, 2]) < 1e-12).astype(np.uint8)
    return c, t, fixed


def _pair(coords, tets, fixed, E=5e5, nu=0.33, rho=1000.0, **kw):
    from opentissue_amd import fem, mesh
    import oracle
    s = fem.Solver(coords, tets, fixed, **kw)
    s.init(E, nu, rho)
    o = oracle.Oracle(coords, tets, fixed)
    o.init(E, nu, rho)
    m = o.mass()
    f = mesh.gravity_forces(np.where(fixed == 1, 0.0, m), fixed)
    s.set_external_forces(f)
    o.set_external_forces(f)
    return s, o


def _per_step_errors(s, o, coords, steps, dt, gparams, oparams, warp=True):
    errs, iters = [], []
    for _ in range(steps):
        xo, vo = o.state()
        s.set_state(xo, vo)
        st = s.simulate(dt, warp, gparams)
        so = o.simulate(dt, warp, oparams)
        xg, _ = s.state()
        xo2, _ = o.state()
        errs.append(rel_disp_error(xg, xo2, coords))
        iters.append((st.iterations, so.iterations))
    return np.array(errs), iters


def test_mass_and_volume_match_oracle():
    from opentissue_amd import fem  # noqa: F401
    c, t, fx = _lattice(3, 4, 2)
    s, o = _pair(c, t, fx)
    m_g, m_o = s.mass(), o.mass()
    free = fx == 0
    assert np.array_equal(m_g[~free], m_o[~free])
    np.testing.assert_allclose(m_g[free], m_o[free], rtol=1e-15, atol=0)
    np.testing.assert_allclose(s.volumes(), o.volumes(), rtol=1e-15, atol=0)


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 3, 2), (10, 40, 10)])
def test_reference_mode_per_step(dims):
    from opentissue_amd import fem
    import oracle
    c, t, fx = _lattice(*dims)
    s, o = _pair(c, t, fx)
    errs, iters = _per_step_errors(s, o, c, 6, 0.01, fem.reference_params(), oracle.reference_params())
    print(dims, errs, iters)
    assert all(i == (20, 20) for i in iters)
    assert errs.max() <= TOL, errs


@pytest.mark.parametrize("dims", [(2, 3, 2), (10, 40, 10)])
def test_config_mode_per_step(dims):
    from opentissue_amd import fem
    import oracle
    c, t, fx = _lattice(*dims)
    s, o = _pair(c, t, fx)
    errs, iters = _per_step_errors(s, o, c, 6, 0.01, fem.config_params(), oracle.config_params())
    print(dims, errs, iters)
    assert errs.max() <= TOL, errs


@pytest.mark.parametrize("block", [256, 512, 1024])
@pytest.mark.parametrize("tile_nodes", [0, 1200])
def test_cg_block_sizes_and_node_batches(block, tile_nodes):
    """Every CG block size, on tiles whose node table fits one register batch
    per thread (default tiling) and on tiles that need several (1,200 owned
    nodes per tile: node tables above 1,024 nodes)."""
    from opentissue_amd import fem
    import oracle
    c, t, fx = _lattice(10, 40, 10)
    s, o = _pair(c, t, fx, target_tile_nodes=tile_nodes)
    s.set_kernel_config(cg_block=block)
    if tile_nodes:
        assert s.layout_info()["max_tile_nodes"] > 1024
    errs, iters = _per_step_errors(s, o, c, 4, 0.01, fem.config_params(), oracle.config_params())
    print(block, tile_nodes, errs, iters)
    assert errs.max() <= TOL, errs


@pytest.mark.parametrize("mode,tol", [("SPMV_F64", TOL), ("SPMV_F64_STORE64", 1e-7)])
def test_spmv_precision_modes(mode, tol):
    """The default SpMV (FP64 arithmetic on FP32 element data) meets the 1e-6
    bar.  With the element data in FP64 too, what is left is summation order:
    the GPU's per-node sums are LDS atomics in wave order, the oracle's are
    sequential, and 130-odd CG iterations carry that difference (measured on
    this mesh: 9e-9 on the first step from rest, ~3e-12 on later steps, at
    identical iteration counts)."""
    from opentissue_amd import fem
    import oracle
    c, t, fx = _lattice(10, 40, 10)
    s, o = _pair(c, t, fx)
    s.set_kernel_config(cg_block=256, spmv_mode=getattr(fem, mode))
    errs, iters = _per_step_errors(s, o, c, 4, 0.01, fem.config_params(), oracle.config_params())
    print(mode, errs, iters)
    assert errs.max() <= tol, errs


@pytest.mark.parametrize("materials", ["uniform", "E per tet", "E and nu per tet"])
def test_slot_record_forms(materials):
    """One Poisson ratio per tile lets the CG's slot records fold E |V| into
    the rotated gradients (44 bytes per slot, D / (E |V|) per tile); a
    Poisson ratio that varies inside a tile keeps D per slot (56 bytes).
    Both match the oracle."""
    from opentissue_amd import fem, mesh
    import oracle
    c, t, fx = _lattice(6, 10, 4)
    rng = np.random.default_rng(3)
    E, nu = np.full(len(t), 3e5), np.full(len(t), 0.3)
    if materials != "uniform":
        E = np.exp(rng.uniform(np.log(1e5), np.log(1e6), len(t)))
    if materials == "E and nu per tet":
        nu = rng.uniform(0.2, 0.45, len(t))
    rho = np.full(len(t), 1000.0)
    s = fem.Solver(c, t, fx)
    s.init_materials(E, nu, rho)
    o = oracle.Oracle(c, t, fx)
    o.init_materials(E, nu, rho)
    f = mesh.gravity_forces(np.where(fx == 1, 0.0, o.mass()), fx)
    s.set_external_forces(f)
    o.set_external_forces(f)
    # CG runs to convergence (heterogeneous E and nu need more than the
    # default cap of 200 iterations; a capped, unconverged step is not
    # comparable at 1e-6), then the reference's fixed 20 iterations
    for gp, op in ((fem.config_params(cap=3000), oracle.config_params(cap=3000)),
                   (fem.reference_params(), oracle.reference_params())):
        errs, iters = _per_step_errors(s, o, c, 3, 0.01, gp, op)
        print(materials, errs, iters)
        assert all(ig < 3000 for ig, _ in iters) or gp.cg_max_iterations == 20
        assert errs.max() <= TOL, errs


@pytest.mark.parametrize("mode", ["ref", "cfg"])
def test_irregular_mesh_read_from_xml(mode):
    """An irregular (TetGen-made) mesh entering through the T4MESH XML reader:
    the reference's fixture demos/data/xml/blocky.xml (438 nodes, 1,831 tets),
    fixed at its lowest y, under gravity along -z.

    The relative-residual mode solves to 1e-8 here.  At 1e-6 this mesh's CG
    converges in ~45 iterations, so two correct runs whose rounding differs
    can stop one iteration apart (measured: 47 against 48 on step 1), and one
    iteration near the stop moves the iterate by ~1e-6 of the first step's
    small displacement (1.1e-6 measured; 1e-8 where the counts agree).  The
    tighter stop compares the arithmetic, not the stopping point."""
    import os
    from opentissue_amd import fem, mesh_io
    import oracle
    c, t = mesh_io.xml_read(os.path.join(os.path.dirname(__file__), "golden", "blocky.xml"))
    fx = (c[:, 1] < c[:, 1].min() + 1e-9).astype(np.uint8)
    assert 0 < fx.sum() < len(fx)
    s, o = _pair(c, t, fx)
    gp = fem.reference_params() if mode == "ref" else fem.config_params(tol=1e-8, cap=3000)
    op = oracle.reference_params() if mode == "ref" else oracle.config_params(tol=1e-8, cap=3000)
    errs, iters = _per_step_errors(s, o, c, 5, 0.01, gp, op)
    print(mode, errs, iters)
    assert errs.max() <= TOL, errs


def test_rotations_match_oracle():
    from opentissue_amd import fem
    import oracle
    c, t, fx = _lattice(4, 6, 3)
    s, o = _pair(c, t, fx)
    P = fem.reference_params()
    P.store_rotations = 1
    for _ in range(3):
        xo, vo = o.state()
        s.set_state(xo, vo)
        s.simulate(0.01, True, P)
        o.simulate(0.01, True, oracle.reference_params())
    # compare the rotation of the last step (both from the same start state)
    Rg, Ro = s.rotations(), o.rotations()
    assert np.abs(Rg - Ro).max() < 1e-12
    assert np.abs(np.einsum("tij,tkj->tik", Rg, Rg) - np.eye(3)).max() < 1e-13


def _bodies(dims_list, seed=0, jitter=0.2):
    from opentissue_amd import mesh
    rng = np.random.default_rng(seed)
    parts, mats, forces, fixed = [], [], [], []
    for b, d in enumerate(dims_list):
        c, t = mesh.generate_blocks(*d, 0.01, 0.01, 0.01)
        if jitter:
            c = mesh.jitter_interior(c, t, 0.01, jitter, seed * 100 + b)
        fx = mesh.face_nodes(c, 0, 0).astype(np.uint8)
        E, nu = float(np.exp(rng.uniform(np.log(1e4), np.log(1e6)))), float(rng.uniform(0.25, 0.45))
        f = mesh.traction_forces(c, t, 0, 1, rng.normal(size=3) * 1e3)
        f[fx == 1] = 0
        parts.append((c, t))
        mats.append((E, nu))
        forces.append(f)
        fixed.append(fx)
    return parts, mats, forces, fixed


@pytest.mark.parametrize("mode", ["ref", "cfg"])
def test_batched_bodies_match_per_body_oracle(mode):
    """Bodies of uneven size stepped together; each has its own CG
    (convergence mask), so each must match an oracle run of that body alone."""
    from opentissue_amd import fem, mesh
    import oracle
    dims = [(2, 3, 2), (3, 3, 3), (1, 2, 4), (5, 2, 2), (1, 1, 1)]
    parts, mats, forces, fixed = _bodies(dims)
    coords, tets, no, to = mesh.pack_bodies(parts)
    E = np.concatenate([np.full(len(t), m[0]) for (c, t), m in zip(parts, mats)])
    nu = np.concatenate([np.full(len(t), m[1]) for (c, t), m in zip(parts, mats)])
    s = fem.Solver(coords, tets, np.concatenate(fixed), body_node_offsets=no, body_tet_offsets=to,
                   target_tile_nodes=16)
    s.init_materials(E, nu, 1000.0)
    s.set_external_forces(np.concatenate(forces))
    gp = fem.reference_params() if mode == "ref" else fem.config_params()
    op = oracle.reference_params() if mode == "ref" else oracle.config_params()
    orcs = []
    for (c, t), (Eb, nub), f, fx in zip(parts, mats, forces, fixed):
        o = oracle.Oracle(c, t, fx)
        o.init(Eb, nub, 1000.0)
        o.set_external_forces(f)
        orcs.append(o)
    for step in range(4):
        xo = np.concatenate([o.state()[0] for o in orcs])
        vo = np.concatenate([o.state()[1] for o in orcs])
        s.set_state(xo, vo)
        s.simulate(0.01, True, gp)
        its_g, _ = s.body_stats()
        xg, _ = s.state()
        for b, o in enumerate(orcs):
            st = o.simulate(0.01, True, op)
            x1, _ = o.state()
            sl = slice(no[b], no[b + 1])
            err = rel_disp_error(xg[sl], x1, parts[b][0])
            assert err <= TOL, (step, b, err)
            # the iteration that crosses rel-res 1e-6 moves with rounding order
            # (LDS-atomic accumulation) and, in this mode, with the FP32
            # storage of the CG direction: by a few iterations (measured: up
            # to 5 of 50 on the smallest bodies, +1% on average)
            slack = max(3, int(0.12 * st.iterations)) if mode == "cfg" else 0
            assert abs(int(its_g[b]) - st.iterations) <= slack, (step, b, its_g[b], st.iterations)


def test_polar_rotation_mode():
    from opentissue_amd import fem
    import oracle
    c, t, fx = _lattice(3, 6, 2)
    s, o = _pair(c, t, fx)
    gp = fem.config_params(rotation_mode=1, store_rotations=1)
    op = oracle.config_params(rotation_mode=1)
    errs, _ = _per_step_errors(s, o, c, 4, 0.01, gp, op)
    assert errs.max() <= TOL, errs
    R = s.rotations()
    assert np.abs(R - o.rotations()).max() < 1e-9
    assert np.all(np.linalg.det(R) > 0)


def test_no_stiffness_warping():
    from opentissue_amd import fem
    import oracle
    c, t, fx = _lattice(3, 6, 2)
    s, o = _pair(c, t, fx)
    errs, _ = _per_step_errors(s, o, c, 4, 0.01, fem.config_params(), oracle.config_params(),
                               warp=False)
    assert errs.max() <= TOL, errs


def test_plasticity_matches_oracle():
    from opentissue_amd import fem
    import oracle
    c, t, fx = _lattice(2, 5, 2)
    s = fem.Solver(c, t, fx)
    o = oracle.Oracle(c, t, fx)
    s.init(5e5, 0.33, 1000.0, 1e-4, 20.0, 0.05)
    o.init(5e5, 0.33, 1000.0, 1e-4, 20.0, 0.05)
    m = o.mass()
    from opentissue_amd import mesh
    f = mesh.gravity_forces(np.where(fx == 1, 0.0, m), fx) * 50.0
    errs = []
    for _ in range(5):
        # the reference accumulates plastic forces into f_external: reset it
        s.set_external_forces(f)
        o.set_external_forces(f)
        xo, vo = o.state()
        s.set_state(xo, vo)
        s.simulate(0.01, True, fem.config_params())
        o.simulate(0.01, True, oracle.config_params())
        errs.append(rel_disp_error(s.state()[0], o.state()[0], c))
        # both must have accumulated the same plastic forces into f_external
        np.testing.assert_allclose(s.external_forces(), o.external_forces(), rtol=1e-6,
                                   atol=1e-9 * np.abs(f).max())
    assert max(errs) <= TOL, errs
    assert np.abs(o.plastic()).max() > 0
    np.testing.assert_allclose(s.plastic(), o.plastic(), rtol=1e-5, atol=1e-9)


def test_fixed_node_velocity_enters_warm_start():
    # r = b - A v uses every column, fixed ones included (fem_conjugate_gradients.h:52-66)
    from opentissue_amd import fem
    import oracle
    c, t, fx = _lattice(2, 4, 2)
    rng = np.random.default_rng(3)
    v0 = 0.05 * rng.normal(size=c.shape)
    s = fem.Solver(c, t, fx, velocity=v0)
    o = oracle.Oracle(c, t, fx, velocity=v0)
    s.init(5e5, 0.33, 1000.0)
    o.init(5e5, 0.33, 1000.0)
    s.simulate(0.01, True, fem.config_params())
    o.simulate(0.01, True, oracle.config_params())
    xg, vg = s.state()
    xo, vo = o.state()
    assert rel_disp_error(xg, xo, c) <= TOL
    # fixed nodes neither move nor change velocity
    np.testing.assert_array_equal(vg[fx == 1], v0[fx == 1])
    np.testing.assert_array_equal(xg[fx == 1], c[fx == 1])


def test_export_K_matches_oracle():
    from opentissue_amd import fem
    import oracle
    c, t, fx = _lattice(2, 3, 2)
    s, o = _pair(c, t, fx)
    P = fem.reference_params()
    P.store_rotations = 1
    for _ in range(3):
        xo, vo = o.state()
        s.set_state(xo, vo)
        s.simulate(0.01, True, P)
        o.simulate(0.01, True, oracle.reference_params())
    rp, cols, blocks = s.export_K()
    N = len(c)
    K = np.zeros((3 * N, 3 * N))
    for i in range(N):
        for k in range(rp[i], rp[i + 1]):
            K[3 * i:3 * i + 3, 3 * cols[k]:3 * cols[k] + 3] = blocks[k]
    Ko = o.export_K_dense()
    np.testing.assert_allclose(K, Ko, rtol=0, atol=1e-9 * np.abs(Ko).max())


def test_update_original_coord_then_reinit():
    from opentissue_amd import fem
    import oracle
    c, t, fx = _lattice(2, 3, 2)
    s, o = _pair(c, t, fx)
    for _ in range(3):
        s.simulate(0.01, True, fem.config_params())
        o.simulate(0.01, True, oracle.config_params())
    s.update_original_coord()
    o.update_original_coord()
    with pytest.raises(fem.OtfemError):
        s.simulate(0.01, True, fem.config_params())  # stale until init (fem_update_original_coord.h:19-23)
    s.init(5e5, 0.33, 1000.0)
    o.init(5e5, 0.33, 1000.0)
    x1, v1 = o.state()
    s.set_state(x1, v1)
    s.simulate(0.01, True, fem.config_params())
    o.simulate(0.01, True, oracle.config_params())
    assert rel_disp_error(s.state()[0], o.state()[0], x1) <= TOL


def test_free_running_drift_relative_residual_mode():
    """The 10x10x40 cantilever (gravity, E 5e5, nu 0.33) free-running 100
    steps on the GPU and in the oracle from the same start, CG to rel-res
    1e-6 with the cap of 200 that this configuration specifies: the
    trajectories stay together (measured: 1.2e-8 at step 1, decaying to
    4e-10; 45 of the steps stop at the cap).  Reference mode is not
    compared free-running: the oracle disagrees with itself there
    (tests/test_oracle.py::test_reference_mode_trajectories_amplify_rounding)."""
    from opentissue_amd import fem, workloads
    import oracle
    w = workloads.cantilever()
    s = fem.Solver(w.coords, w.tets, w.fixed)
    s.init_materials(w.young, w.poisson, w.density)
    s.set_external_forces(w.forces)
    o = oracle.Oracle(w.coords, w.tets, w.fixed)
    o.init_materials(w.young, w.poisson, w.density)
    o.set_external_forces(w.forces)
    gp, op = fem.config_params(cap=200), oracle.config_params(cap=200)
    errs = []
    for _ in range(100):
        s.simulate(w.dt, True, gp)
        o.simulate(w.dt, True, op)
        errs.append(rel_disp_error(s.state()[0], o.state()[0], w.coords))
    print("free-running drift: max", max(errs), "last", errs[-1], flush=True)
    assert max(errs) <= TOL


@pytest.mark.parametrize("rotation_mode", [0, 1], ids=["gram_schmidt", "polar"])
def test_rotations_of_random_deformations(rotation_mode):
    """The pattern of the reference's unit_poloar_decomposition.cpp:26-60 on
    the GPU path: 20,000 disconnected unit tets, each deformed by
    A = R0 (S^T S + 0.1 I) with R0 a random proper rotation, rotated out by
    one step with stored rotations.  Both modes: orthonormal, det +1, equal
    to the oracle's (the reference's 1e-6).  Polar mode, the closed form:
    the polar factor of A is R0, and R^T A is symmetric."""
    from opentissue_amd import fem
    import oracle
    rng = np.random.default_rng(11)
    T, h = 20000, 0.01
    rest = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]]) * h
    x0 = np.tile(rest, (T, 1))
    tets = np.arange(4 * T, dtype=np.int32).reshape(T, 4)
    Q, Rq = np.linalg.qr(rng.normal(size=(T, 3, 3)))
    Q = Q * np.sign(np.diagonal(Rq, axis1=1, axis2=2))[:, None, :]
    Q[np.linalg.det(Q) < 0, :, 0] *= -1
    S = rng.normal(size=(T, 3, 3))
    A = Q @ (S.transpose(0, 2, 1) @ S + 0.1 * np.eye(3))
    x = (np.einsum("tij,kj->tki", A, rest) + rng.normal(size=(T, 1, 3)) * h).reshape(-1, 3)
    s = fem.Solver(x0, tets)
    s.init(5e5, 0.33, 1000.0)
    s.set_state(x, np.zeros_like(x))
    P = fem.reference_params()
    P.rotation_mode, P.store_rotations = rotation_mode, 1
    s.simulate(0.001, True, P)
    R = s.rotations()
    o = oracle.Oracle(x0, tets)
    o.init(5e5, 0.33, 1000.0)
    o.set_state(x, np.zeros_like(x))
    op = oracle.reference_params()
    op.rotation_mode = rotation_mode
    o.simulate(0.001, True, op)
    Ro = o.rotations()
    orth = np.abs(np.einsum("tij,tkj->tik", R, R) - np.eye(3)).max()
    vs_oracle = np.abs(R - Ro).max()
    print("orthonormality", orth, "vs oracle", vs_oracle, flush=True)
    assert orth < 1e-12
    assert (np.linalg.det(R) > 0).all()
    assert vs_oracle < 1e-6
    if rotation_mode == 1:
        RtA = R.transpose(0, 2, 1) @ A
        closed = np.abs(R - Q).max()
        sym = (np.abs(RtA - RtA.transpose(0, 2, 1)).max(axis=(1, 2)) / np.abs(A).max(axis=(1, 2))).max()
        print("vs closed form", closed, "asymmetry of R^T A", sym, flush=True)
        assert closed < 1e-9
        assert sym < 1e-6


@pytest.mark.parametrize("mode", ["SPMV_F64_STORE64", "SPMV_F64"])
def test_stop_measures_the_true_residual(mode):
    """CG stops at |r| <= 1e-6 |b| (the relative-residual extension of
    fem_conjugate_gradients.h:110-119), r being the residual of the GPU's own
    operator.  Checked against the system the oracle assembles for the same
    step from the same state (A = (1 + alpha dt) M + (dt^2 + beta dt) K' and
    b, free rows, all columns), on a stiff jittered body under a traction
    (about 200 CG iterations per step), five steps from shared states:
    * FP64 element records (SPMV_F64_STORE64): the GPU's operator is the
      oracle's to rounding, and |b - A v_gpu| / |b| <= 1e-6 (measured equal
      to the GPU's own recursive residual to 4 digits);
    * the default (SPMV_F64: FP32 slot records, FP32 direction p, and r and
      A p in FP32 once the relative residual is below 1e-2):
      |b - A v_gpu| / |b| <= 2.5e-6.  The excess over 1e-6 is the FP32
      rounding of the element records, whose node sums cancel (measured
      0.97-2.1e-6, the same with FP64 directions), not the vectors'
      storage: p is rounded before every use, so r stays the residual of v,
      and the rounding of r and A p below the switch moves r from b - A v by
      about 1% of the tolerance.
    The displacement agrees with the oracle's converged step to 1e-6."""
    from opentissue_amd import fem, mesh
    import oracle
    h = 0.01
    c0, t = mesh.generate_blocks(8, 5, 4, h, h, h)
    c = mesh.jitter_interior(c0, t, h, 0.25, 7)
    fx = mesh.face_nodes(c, 0, 0).astype(np.uint8)
    rng = np.random.default_rng(3)
    E = np.exp(rng.uniform(np.log(1e5), np.log(1e6), size=len(t)))
    nu = np.full(len(t), 0.3)
    f = mesh.traction_forces(c, t, 0, 1, np.array([0.0, 400.0, -1500.0]))
    f[fx == 1] = 0.0
    s = fem.Solver(c, t, fx)  # 160 cubes x 5 = 800 tets, 270 nodes
    s.init_materials(E, nu, 1000.0)
    s.set_kernel_config(spmv_mode=getattr(fem, mode))
    o = oracle.Oracle(c, t, fx)
    o.init_materials(E, nu, 1000.0)
    s.set_external_forces(f)
    o.set_external_forces(f)
    gp, op = fem.config_params(cap=5000), oracle.config_params(cap=5000)
    free = np.repeat(fx == 0, 3)
    bound = 1e-6 if mode == "SPMV_F64_STORE64" else 2.5e-6
    for step in range(5):
        xo, vo = o.state()
        s.set_state(xo, vo)
        st = s.simulate(0.01, True, gp)
        xg, vg = s.state()
        o.simulate(0.01, True, op)
        # free rows only: fixed rows carry the reference's DBL_MAX mass
        A, b = o.dense(1)[free], o.rhs().reshape(-1)[free]
        res = np.linalg.norm(b - A @ vg.reshape(-1)) / np.linalg.norm(b)
        x1, _ = o.state()
        err = rel_disp_error(xg, x1, xo)
        print(mode, step, st.iterations, res, np.sqrt(st.rr / st.bb), err)
        assert st.iterations > 100 and st.bodies_converged == 1, (step, st.iterations)
        assert res <= bound and err <= TOL, (step, res, err)


def test_modes_alternate_on_one_solver():
    """The CG direction's storage follows the step's settings (FP32 with a
    relative-residual stop, FP64 in the reference mode's fixed 20
    iterations), in one buffer of the solver: steps that alternate the two
    modes on one solver each match the oracle from a shared state (per-step
    displacement error <= 1e-6; the relative-residual steps solved to 1e-8)."""
    from opentissue_amd import fem
    import oracle
    c, t, fx = _lattice(4, 10, 3)
    s, o = _pair(c, t, fx)
    modes = [(fem.reference_params(), oracle.reference_params()),
             (fem.config_params(tol=1e-8, cap=3000), oracle.config_params(tol=1e-8, cap=3000))]
    for step in range(6):
        gp, op = modes[step % 2]
        errs, iters = _per_step_errors(s, o, c, 1, 0.01, gp, op)
        print(step, errs, iters)
        assert errs.max() <= TOL, (step, errs, iters)


def test_capped_steps_with_more_tiles_than_resident_workgroups():
    """Reference-mode steps end at the iteration cap, so every body
    completes its last update in k_finish.  With 1024-thread tiles (one
    workgroup per CU) and 621 tiles, workgroups of the body start after its
    first tile has recorded the stop; they still complete their nodes (the
    stop is recorded as that of kernel cg_max), and each step matches the
    oracle from a shared state."""
    from opentissue_amd import fem
    import oracle
    c, t, fx = _lattice(10, 40, 10)
    s, o = _pair(c, t, fx, target_tile_nodes=8)
    s.set_kernel_config(cg_block=1024)
    assert s.layout_info()["tiles"] > 256
    errs, iters = _per_step_errors(s, o, c, 3, 0.01, fem.reference_params(), oracle.reference_params())
    print(s.layout_info()["tiles"], errs, iters)
    assert errs.max() <= TOL, errs
