"""GPU: the subspace CG start (otfem_params.subspace).

CG starts from v^n + D c, D = [v^n, differences of the last velocities
(, the network's start - v^n)], c the A-optimal coefficients per body; CG
then runs to the same stopping rule.  Checked here:

* per-step parity with the FP64 oracle (plain start) from the GPU's state,
  both solved to rel-res 1e-9 (a converged solve does not depend on the
  start), <= 1e-6 of the step's displacement;
* it cuts CG iterations against the plain start on a cantilever episode;
* bodies stepped in a batch and alone follow the same trajectories;
* with a network attached, its start is one more direction and the result
  is again the converged solve;
* the deterministic mode gives the same bits on every run.
"""
import numpy as np
import pytest

from conftest import rel_disp_error

pytestmark = pytest.mark.gpu


def _cantilever_solver(**kw):
    from opentissue_amd import warmstart, workloads
    w = workloads.cantilever(**kw)
    s = warmstart.make_solver(w)
    return w, s


def test_subspace_per_step_parity_with_oracle():
    from opentissue_amd import fem, warmstart
    import oracle
    w, s = _cantilever_solver()
    load = warmstart.random_loads(w, 1, seed=7)[0][0]
    s.set_external_forces(load)
    P = fem.config_params(tol=1e-9, cap=5000, subspace=6)
    for _ in range(10):  # build the velocity history
        s.simulate(w.dt, True, P)
    o = oracle.Oracle(w.coords, w.tets, w.fixed)
    o.init_materials(w.young, w.poisson, w.density)
    o.set_external_forces(load)
    PO = oracle.config_params(tol=1e-9, cap=5000)
    errs = []
    for _ in range(5):
        xg, vg = s.state()
        o.set_state(xg, vg)
        st = s.simulate(w.dt, True, P)
        so = o.simulate(w.dt, True, PO)
        x1, _ = s.state()
        xo, _ = o.state()
        errs.append(rel_disp_error(x1, xo, xg))
        print("GPU subspace", st.iterations, "oracle plain", so.iterations, "error", errs[-1])
        assert st.bodies_converged == 1
    assert max(errs) <= 1e-6, errs


def _galerkin_start(A, b, v, dirs, thr=1e-12):
    """The A-optimal start v + D c over the directions (columns of D) with
    the same unit-diagonal scaling and pivot dropping as the device."""
    D = np.stack(dirs, 1)
    r = b - A @ v
    G = D.T @ (A @ D)
    g = D.T @ r
    k = len(g)
    s = np.where(np.diag(G) > 0, 1 / np.sqrt(np.maximum(np.diag(G), 1e-300)), 0.0)
    Gs = G * s[:, None] * s[None, :]
    keep = []
    for j in range(k):
        if s[j] == 0:
            continue
        piv = Gs[j, j] - (0 if not keep else np.linalg.solve(Gs[np.ix_(keep, keep)], Gs[keep, j]) @ Gs[keep, j])
        if piv > thr:
            keep.append(j)
    c = np.zeros(k)
    c[keep] = np.linalg.solve(Gs[np.ix_(keep, keep)], (g * s)[keep]) * s[keep]
    w = D @ c
    return v + w, r - A @ w, c


def test_subspace_start_matches_cpu_galerkin():
    """The device start's residual equals the Galerkin start computed on the
    CPU from the oracle's assembled A and b at the same state and the same
    (FP32-rounded) velocity differences."""
    import scipy.sparse as sp
    from opentissue_amd import fem, warmstart
    import oracle
    w, s = _cantilever_solver(I=4, J=12, K=4)
    load = warmstart.random_loads(w, 1, seed=11)[0][0]
    s.set_external_forces(load)
    H = 6
    P = fem.config_params(tol=1e-10, cap=5000, subspace=H)
    vs = []
    for _ in range(10):
        vs.append(s.state()[1].ravel().copy())
        s.simulate(w.dt, True, P)
    free = np.repeat(w.fixed == 0, 3)
    o = oracle.Oracle(w.coords, w.tets, w.fixed)
    o.init_materials(w.young, w.poisson, w.density)
    o.set_external_forces(load)
    N = w.num_nodes
    for _ in range(3):
        x, v = s.state()
        o.set_state(x, v)
        o.simulate(w.dt, True, oracle.config_params(cap=0))
        rp, cols, blocks = o.rows(1)
        A = sp.bsr_matrix((blocks, cols, rp), shape=(3 * N, 3 * N)).tocsr()[free][:, free]
        b = o.rhs().ravel()[free]
        hist = [v.ravel()] + vs[::-1]  # v^n, v^n-1, ...
        dirs = [hist[0][free]]
        for j in range(1, H + 1):
            dirs.append((hist[j - 1] - hist[j]).astype(np.float32).astype(np.float64)[free])
        _, r0, c_cpu = _galerkin_start(A, b, hist[0][free], dirs)
        r_plain = b - A @ hist[0][free]
        st = s.simulate(w.dt, True, P)
        c_gpu = s.subspace_coefficients()[0][:H + 1]
        res_gpu = np.sqrt(st.rr0)
        print("|r0| device", res_gpu, "cpu", np.linalg.norm(r0), "plain", np.linalg.norm(r_plain),
              "\nc device", c_gpu, "\nc cpu   ", c_cpu)
        assert abs(res_gpu - np.linalg.norm(r0)) <= 2e-2 * np.linalg.norm(r0)
        assert res_gpu < 0.2 * np.linalg.norm(r_plain)
        vs.append(v.ravel().copy())


def test_subspace_cuts_iterations():
    from opentissue_amd import fem, warmstart
    w, s = _cantilever_solver()
    loads, _ = warmstart.random_loads(w, 1, seed=1001)
    its = {}
    for k in (0, 3, 6, 10):
        P = fem.config_params(cap=3000, subspace=k)
        s.set_state(w.coords, np.zeros_like(w.coords))
        s.set_external_forces(loads[0])
        seq = []
        for _ in range(40):
            st = s.simulate(w.dt, True, P)
            assert st.bodies_converged == 1
            seq.append(st.iterations)
        its[k] = np.array(seq)
        print("subspace", k, "iterations per step", seq)
    assert its[10][10:].mean() < 0.5 * its[0][10:].mean()
    assert its[6][10:].mean() < 0.7 * its[0][10:].mean()
    assert its[3][10:].mean() < its[0][10:].mean()


def test_subspace_batch_equals_bodies_alone():
    from opentissue_amd import fem, workloads
    w = workloads.bodies(6, 3, 7, seed=3)
    P = fem.config_params(tol=1e-10, cap=5000, subspace=5)
    sb = fem.Solver(w.coords, w.tets, w.fixed, body_node_offsets=w.body_node_offsets,
                    body_tet_offsets=w.body_tet_offsets)
    sb.init_materials(w.young, w.poisson, w.density)
    sb.set_external_forces(w.forces)
    for _ in range(12):
        st = sb.simulate(w.dt, True, P)
        assert st.bodies_converged == w.num_bodies
    xb, _ = sb.state()
    no, to = w.body_node_offsets, w.body_tet_offsets
    for b in range(w.num_bodies):
        ns, ts = slice(no[b], no[b + 1]), slice(to[b], to[b + 1])
        s = fem.Solver(w.coords[ns], w.tets[ts] - no[b], w.fixed[ns])
        s.init_materials(w.young[ts], w.poisson[ts], w.density[ts])
        s.set_external_forces(w.forces[ns])
        for _ in range(12):
            s.simulate(w.dt, True, P)
        x, _ = s.state()
        err = rel_disp_error(xb[ns], x, w.coords[ns])
        print("body", b, "batch vs alone", err)
        assert err <= 1e-8


def test_subspace_with_network_direction():
    from opentissue_amd import fem, warmstart
    w, s = _cantilever_solver()
    load = warmstart.random_loads(w, 1, seed=5)[0][0]
    model = fem.Model(seed=0)  # quadratic extrapolation of the velocity
    s.set_warmstart_model(model)
    runs = {}
    for learned in (False, True):
        P = fem.config_params(tol=1e-9, cap=5000, subspace=4)
        P.warm_start = fem.WARM_LEARNED if learned else fem.WARM_PREVIOUS_VELOCITY
        s.set_state(w.coords, np.zeros_like(w.coords))
        s.set_external_forces(load)
        its = []
        for _ in range(20):
            st = s.simulate(w.dt, True, P)
            assert st.bodies_converged == 1 and st.flags == 0
            its.append(st.iterations)
        runs[learned] = (s.state()[0], np.array(its))
        print("network direction" if learned else "history only", its)
    err = rel_disp_error(runs[True][0], runs[False][0], w.coords)
    print("trajectory difference", err)
    assert err <= 1e-7
    # one more direction can only lower the start's A-norm error
    assert runs[True][1][8:].mean() <= 1.05 * runs[False][1][8:].mean() + 1


def test_subspace_deterministic_bits():
    from opentissue_amd import fem, warmstart
    w, s = _cantilever_solver(I=4, J=10, K=4)
    load = warmstart.random_loads(w, 1, seed=3)[0][0]
    s.set_deterministic(True)
    P = fem.config_params(cap=3000, subspace=5)
    out = []
    for _ in range(2):
        s.set_state(w.coords, np.zeros_like(w.coords))
        s.set_external_forces(load)
        for _ in range(15):
            s.simulate(w.dt, True, P)
        out.append(s.state())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_subspace_lds_limit_is_an_error_not_a_launch():
    """A tiling whose tile node table with the requested directions exceeds
    the workgroup's LDS (dynamic plus static) is rejected by the step with
    OTFEM_INVALID_ARG before any launch; fewer directions still run."""
    from opentissue_amd import fem, workloads
    w = workloads.cantilever()
    s = fem.Solver(w.coords, w.tets, w.fixed, target_tile_nodes=1500)
    s.init_materials(w.young, w.poisson, w.density)
    s.set_external_forces(w.forces)
    assert s.layout_info()["max_tile_nodes"] > 1200
    model = fem.Model(seed=0)
    s.set_warmstart_model(model)
    P = fem.config_params(cap=3000, subspace=10)
    P.warm_start = fem.WARM_LEARNED
    with pytest.raises(fem.OtfemError) as e:
        s.simulate(w.dt, True, P)
    assert e.value.status == fem.INVALID_ARG
    P1 = fem.config_params(cap=3000, subspace=1)
    for _ in range(3):
        st = s.simulate(w.dt, True, P1)
        assert st.bodies_converged == 1


def test_subspace_learned_on_the_64_cube_tiles():
    """The learned start with the most directions (10 differences and the
    network's) on the 64^3 cube's default tiles, whose largest node table
    (~980 nodes) is the tightest fit for the Gram kernel's LDS."""
    from opentissue_amd import fem, workloads
    w = workloads.cube64()
    s = fem.Solver(w.coords, w.tets, w.fixed)
    s.init_materials(w.young, w.poisson, w.density)
    s.set_external_forces(w.forces)
    model = fem.Model(seed=0)
    s.set_warmstart_model(model)
    P = fem.config_params(cap=5000, subspace=10)
    P.warm_start = fem.WARM_LEARNED
    its = []
    for _ in range(14):
        st = s.simulate(w.dt, True, P)
        assert st.bodies_converged == 1 and st.flags == 0
        its.append(st.iterations)
    print("64^3 cube, learned start with 12 directions:", its)
    assert np.mean(its[-4:]) < 0.8 * its[0]
