"""Learned CG warm start: trajectories, training, learned-vs-plain evaluation.

The network and its kernels are in ``csrc/mlp.hip`` (``fem.Model``).  This
module drives them the way the learned warm start is meant to be built:

1. run the *plain* solver (CG started from v^n, the reference's start,
   fem_conjugate_gradients.h:48-67) on episodes that start at rest under a
   seeded random traction (0.5e3 .. 2e3 Pa on a random free face, random
   direction) plus the workload's own loads, recording one sample per free
   node and step;
2. train the network on those samples with Adam (device kernels);
3. evaluate on held-out seeds: the same episodes stepped with the plain start
   and with the learned start, CG to the same stopping rule in both.
"""
from __future__ import annotations

import time

import numpy as np

from . import fem, mesh


def free_faces(coords, fixed):
    """Bounding-box faces (axis, side) that carry no fixed node."""
    out = []
    for axis in range(3):
        for side in range(2):
            on = mesh.face_nodes(coords, axis, side)
            if on.any() and not np.any(fixed[on] == 1):
                out.append((axis, side))
    return out


def random_loads(workload, count, seed, magnitude=(0.5e3, 2e3)):
    """`count` force fields: the workload's forces plus a traction of random
    magnitude and direction on a random free face.  Returns (forces, meta)."""
    rng = np.random.default_rng(seed)
    c, t, fx = workload.coords, workload.tets, workload.fixed
    faces = mesh.boundary_faces(t)
    choices = free_faces(c, fx)
    loads, meta = [], []
    for _ in range(count):
        axis, side = choices[rng.integers(len(choices))]
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        tr = rng.uniform(*magnitude) * d
        f = workload.forces + mesh.traction_forces(c, t, axis, side, tr, faces)
        f[fx == 1] = 0.0
        loads.append(f)
        meta.append(dict(axis=int(axis), side=int(side), traction=tr.tolist()))
    return loads, meta


def make_solver(workload, **kw):
    s = fem.Solver(workload.coords, workload.tets, workload.fixed,
                   body_node_offsets=workload.body_node_offsets if workload.num_bodies > 1 else None,
                   body_tet_offsets=workload.body_tet_offsets if workload.num_bodies > 1 else None,
                   **kw)
    s.init_materials(workload.young, workload.poisson, workload.density)
    return s


def _rest(solver, workload):
    solver.set_state(workload.coords, np.zeros_like(workload.coords))


def record(solver, model, workload, loads, steps, params, dt=None):
    """Plain-solver trajectories from rest under each load; every step appends
    one sample per free node to the model's store."""
    dt = workload.dt if dt is None else dt
    P = fem.Params.from_buffer_copy(params)
    P.warm_start = fem.WARM_PREVIOUS_VELOCITY
    n_free = int((workload.fixed == 0).sum())
    model.reserve(model.num_samples() + len(loads) * steps * n_free)
    solver.set_warmstart_model(model)
    solver.set_recording(True)
    try:
        for f in loads:
            _rest(solver, workload)
            solver.set_external_forces(f)
            for _ in range(steps):
                solver.step_async(dt, True, P)
            solver.synchronize()
    finally:
        solver.set_recording(False)
    return model.num_samples()


def run_episodes(solver, workload, loads, steps, params, learned, model=None, dt=None):
    """Step every load's episode from rest; returns per-step CG iterations
    (body maximum), total wall seconds of the stepping and device-side stats."""
    dt = workload.dt if dt is None else dt
    P = fem.Params.from_buffer_copy(params)
    P.warm_start = fem.WARM_LEARNED if learned else fem.WARM_PREVIOUS_VELOCITY
    solver.set_warmstart_model(model if learned else None)
    iters, secs = [], 0.0
    for f in loads:
        _rest(solver, workload)
        solver.set_external_forces(f)
        solver.iteration_totals(reset=True)
        t0 = time.perf_counter()
        for _ in range(steps):
            solver.step_async(dt, True, P)
        st = solver.synchronize()
        secs += time.perf_counter() - t0
        if st.flags & 1:
            raise fem.OtfemError(fem.NONFINITE, "non-finite CG scalar")
        iters.append(solver.iteration_totals().astype(np.float64) / steps)
    T = workload.num_tets
    return dict(mean_iterations=float(np.mean([i.mean() for i in iters])),
                seconds=secs, steps=steps * len(loads),
                tet_steps_per_s=T * steps * len(loads) / secs)


def per_step_iterations(solver, workload, load, steps, params, learned, model=None, dt=None):
    """Synchronous version for small cases: per-step iterations and r0/b."""
    dt = workload.dt if dt is None else dt
    P = fem.Params.from_buffer_copy(params)
    P.warm_start = fem.WARM_LEARNED if learned else fem.WARM_PREVIOUS_VELOCITY
    solver.set_warmstart_model(model if learned else None)
    _rest(solver, workload)
    solver.set_external_forces(load)
    its, rel0 = [], []
    for _ in range(steps):
        st = solver.simulate(dt, True, P)
        its.append(st.iterations)
        rel0.append(np.sqrt(st.rr0 / st.bb) if st.bb > 0 else 0.0)
    return np.array(its), np.array(rel0)
