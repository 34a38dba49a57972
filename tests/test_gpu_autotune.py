"""Autotuners: otfem_autotune (CG block size on a live solver) and
otfem_autotune_layout (tile target and block size on throw-away solvers).

They choose launch shapes only, so the checks are that the choice is one of
the candidates, that tuning leaves the solver's state as it was, that a layout
beyond the LDS budget is skipped, and that a solver built with the chosen
layout steps to the same result as the default layout within the parity bar
(1e-6 of the displacement).  They differ by rounding only: both run CG to
rel-res 1e-6, with per-node sums in another order when the tile or block size
differs (measured 2e-8 after three steps from rest on a 30^3 cube with
blocks of 512 and 256 threads).
"""
import numpy as np
import pytest

from opentissue_amd import fem, mesh

pytestmark = pytest.mark.gpu

H = 0.01


def _cube(n, jitter=0.2, seed=0):
    c0, t = mesh.generate_blocks(n, n, n, H, H, H)
    c = mesh.jitter_interior(c0, t, H, jitter, seed)
    fixed = mesh.face_nodes(c, 0, 0).astype(np.uint8)
    f = mesh.traction_forces(c, t, 0, 1, np.array([0.0, 0.0, -1e3]))
    f[fixed == 1] = 0.0
    E = np.exp(np.random.default_rng(seed + 1).uniform(np.log(1e5), np.log(1e6), size=len(t)))
    return c, t, fixed, f, E


def _run(c, t, fixed, f, E, P, steps, tile=0, block=0):
    s = fem.Solver(c, t, fixed, target_tile_nodes=tile)
    s.init_materials(E, np.full(len(t), 0.3), np.full(len(t), 1000.0))
    s.set_external_forces(f)
    if block:
        s.set_kernel_config(cg_block=block)
    for _ in range(steps):
        s.simulate(0.005, True, P)
    x, v = s.state()
    s.close()
    return x


def test_block_autotune_restores_state():
    """Node state, rotations, per-body iteration / unconverged totals, the
    last step's body records and the timing totals are the same before and
    after the trials (which run steps of their own)."""
    c, t, fixed, f, E = _cube(12)
    P = fem.config_params(store_rotations=1)
    s = fem.Solver(c, t, fixed)
    s.init_materials(E, np.full(len(t), 0.3), np.full(len(t), 1000.0))
    s.set_external_forces(f)
    s.set_timing(True)
    s.simulate(0.005, True, P)
    s.step_async(0.005, True, P)  # pending when the autotuner starts
    s.synchronize()
    x0, v0 = s.state()
    R0 = s.rotations()
    it0, un0 = s.iteration_totals(), s.unconverged_totals()
    bs0 = s.body_stats()
    tm0 = s.timing_detail()
    block = s.autotune(0.005, True, P, steps=2)
    print("block", block, flush=True)
    assert block in (256, 512, 1024)
    x1, v1 = s.state()
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1)
    assert np.array_equal(s.external_forces(), f)
    assert np.array_equal(s.rotations(), R0) and not np.allclose(R0, np.eye(3))
    assert np.array_equal(s.iteration_totals(), it0) and np.array_equal(s.unconverged_totals(), un0)
    bs1 = s.body_stats()
    assert all(np.array_equal(a, b) for a, b in zip(bs0, bs1))
    assert s.timing_detail() == tm0
    s.close()


def test_layout_autotune_choice_steps_like_default():
    c, t, fixed, f, E = _cube(30)
    P = fem.config_params(cap=5000)  # converged steps: capped ones amplify rounding
    nu, rho = np.full(len(t), 0.3), np.full(len(t), 1000.0)
    # 20,000 owned nodes per tile cannot fit the LDS: skipped
    tile, block, ms = fem.autotune_layout(c, t, fixed, E, nu, rho, f, 0.005, P, steps=2,
                                          candidates=[0, 2000, 20000])
    print("tile", tile, "block", block, "ms", ms, flush=True)
    assert tile in (0, 2000) and block in (256, 512, 1024) and ms > 0
    x_tuned = _run(c, t, fixed, f, E, P, 3, tile, block)
    x_def = _run(c, t, fixed, f, E, P, 3)
    disp = np.abs(x_def - c).max()
    err = np.abs(x_tuned - x_def).max() / disp
    print("relative difference", err, flush=True)
    assert err <= 1e-6


def test_layout_autotune_default_candidates_on_bodies():
    rng = np.random.default_rng(2)
    dims = rng.integers(3, 12, size=(20, 3))
    parts = [mesh.generate_blocks(int(I), int(J), int(K), H, H, H) for I, J, K in dims]
    c, t, no, to = mesh.pack_bodies(parts)
    fixed = np.zeros(len(c), np.uint8)
    for b in range(len(dims)):
        fixed[no[b]:no[b + 1]] = mesh.face_nodes(c[no[b]:no[b + 1]], 0, 0)
    tile, block, ms = fem.autotune_layout(c, t, fixed, 1e5, 0.3, 1000.0, None, 0.01,
                                          fem.config_params(), steps=1,
                                          body_node_offsets=no, body_tet_offsets=to)
    print("tile", tile, "block", block, flush=True)
    assert tile in (256, 512, 1024) and block in (256, 512, 1024)


def test_layout_autotune_reports_when_nothing_fits():
    c, t, fixed, f, E = _cube(30)
    with pytest.raises(fem.OtfemError, match="no candidate layout could be stepped"):
        fem.autotune_layout(c, t, fixed, E, 0.3, 1000.0, f, 0.005, fem.config_params(),
                            candidates=[20000])
