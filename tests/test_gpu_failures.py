"""GPU: every forceable failure yields a defined status, not a crash or a
silent NaN (inverted / zero-volume tets, materials out of range, free nodes
outside every tet, non-finite forces, an all-fixed body, a cap of one CG
iteration, a mesh with no tets).  All conditions are set from the host."""
import numpy as np
import pytest

from conftest import rel_disp_error

pytestmark = pytest.mark.gpu


def _lattice(I=2, J=3, K=2):
    from opentissue_amd import mesh
    c, t = mesh.generate_blocks(I, J, K, 0.01, 0.01, 0.01)
    fx = (np.abs(c[:, 2]) < 1e-12).astype(np.uint8)
    return c, t, fx


def _status(err):
    return err.value.status


def test_inverted_and_degenerate_tets_are_rejected():
    from opentissue_amd import fem
    c, t, fx = _lattice()
    bad = t.copy()
    bad[3, [1, 2]] = bad[3, [2, 1]]  # orientation flipped: V < 0
    s = fem.Solver(c, bad, fx)
    with pytest.raises(fem.OtfemError) as e:
        s.init(5e5, 0.33, 1000.0)
    assert _status(e) == fem.INVERTED_TET
    flat = c.copy()
    flat[t[0, 3]] = flat[t[0, 0]] + (flat[t[0, 1]] - flat[t[0, 0]]) * 0.5  # corner on an edge
    s = fem.Solver(flat, t, fx)
    with pytest.raises(fem.OtfemError) as e:
        s.init(5e5, 0.33, 1000.0)
    assert _status(e) == fem.INVERTED_TET


@pytest.mark.parametrize("E,nu,rho", [(5e5, 0.5, 1000.0), (5e5, 0.7, 1000.0), (5e5, -0.1, 1000.0),
                                      (-1.0, 0.3, 1000.0), (5e5, 0.3, 0.0), (np.inf, 0.3, 1000.0)])
def test_materials_out_of_range_are_rejected(E, nu, rho):
    from opentissue_amd import fem
    c, t, fx = _lattice()
    s = fem.Solver(c, t, fx)
    with pytest.raises(fem.OtfemError) as e:
        s.init(E, nu, rho)
    assert _status(e) == fem.INVALID_ARG
    with pytest.raises(fem.OtfemError) as e:
        s.simulate(0.01)  # still not initialised
    assert _status(e) == fem.STATE


def test_free_node_outside_every_tet_is_rejected():
    from opentissue_amd import fem
    c, t, fx = _lattice()
    c2 = np.vstack([c, [[1.0, 1.0, 1.0]]])
    s = fem.Solver(c2, t, np.append(fx, 0))
    with pytest.raises(fem.OtfemError) as e:
        s.init(5e5, 0.33, 1000.0)
    assert _status(e) == fem.INVALID_ARG
    s = fem.Solver(c2, t, np.append(fx, 1))  # fixed: allowed
    s.init(5e5, 0.33, 1000.0)
    s.simulate(0.01, True, fem.config_params())


def test_nonfinite_force_reports_nonfinite_and_recovers():
    from opentissue_amd import fem, mesh
    c, t, fx = _lattice()
    s = fem.Solver(c, t, fx)
    s.init(5e5, 0.33, 1000.0)
    f = mesh.gravity_forces(np.where(fx == 1, 0.0, s.mass()), fx)
    bad = f.copy()
    bad[np.nonzero(fx == 0)[0][0], 2] = np.nan
    s.set_external_forces(bad)
    with pytest.raises(fem.OtfemError) as e:
        s.simulate(0.01, True, fem.config_params())
    assert _status(e) == fem.NONFINITE
    s.set_state(c, np.zeros_like(c))
    s.set_external_forces(f)
    st = s.simulate(0.01, True, fem.config_params())
    assert st.flags == 0 and st.bodies_converged == 1
    assert np.all(np.isfinite(s.state()[0]))


def test_all_fixed_body_is_a_no_op_beside_a_moving_one():
    from opentissue_amd import fem, mesh
    import oracle
    c1, t1, f1 = _lattice()
    c2, t2, _ = _lattice(1, 2, 1)
    c2 = c2 + np.array([1.0, 0.0, 0.0])
    coords, tets, no, to = mesh.pack_bodies([(c1, t1), (c2, t2)])
    fixed = np.concatenate([f1, np.ones(len(c2), dtype=np.uint8)])
    s = fem.Solver(coords, tets, fixed, body_node_offsets=no, body_tet_offsets=to)
    s.init(5e5, 0.33, 1000.0)
    m = s.mass()
    f = mesh.gravity_forces(np.where(fixed == 1, 0.0, m), fixed)
    s.set_external_forces(f)
    o = oracle.Oracle(c1, t1, f1)
    o.init(5e5, 0.33, 1000.0)
    o.set_external_forces(f[:len(c1)])
    for _ in range(3):
        st = s.simulate(0.01, True, fem.config_params())
        o.simulate(0.01, True, oracle.config_params())
        assert st.flags == 0 and st.bodies_converged == 2
    its, _ = s.body_stats()
    assert its[1] == 0
    x, v = s.state()
    assert np.array_equal(x[len(c1):], c2) and np.all(v[len(c1):] == 0)
    assert rel_disp_error(x[:len(c1)], o.state()[0], c1) <= 1e-6


def test_cap_of_one_iteration():
    from opentissue_amd import fem, mesh
    import oracle
    c, t, fx = _lattice()
    s = fem.Solver(c, t, fx)
    s.init(5e5, 0.33, 1000.0)
    o = oracle.Oracle(c, t, fx)
    o.init(5e5, 0.33, 1000.0)
    f = mesh.gravity_forces(np.where(fx == 1, 0.0, o.mass()), fx)
    s.set_external_forces(f)
    o.set_external_forces(f)
    s.unconverged_totals(reset=True)
    st = s.simulate(0.01, True, fem.config_params(cap=1))
    so = o.simulate(0.01, True, oracle.config_params(cap=1))
    assert st.iterations == 1 == so.iterations and st.bodies_capped == 1
    assert int(s.unconverged_totals()[0]) == 1
    assert rel_disp_error(s.state()[0], o.state()[0], c) <= 1e-6


def test_mesh_without_tets_and_all_nodes_fixed():
    from opentissue_amd import fem
    c = np.random.default_rng(0).normal(size=(5, 3))
    s = fem.Solver(c, np.zeros((0, 4), dtype=np.int32), np.ones(5, dtype=np.uint8))
    s.init(5e5, 0.33, 1000.0)
    st = s.simulate(0.01, True, fem.config_params())
    assert st.flags == 0 and st.iterations == 0
    assert np.array_equal(s.state()[0], c)
