"""GPU: what the trained learned warm start gains over the plain start.

Protocol (the 10x10x40 cantilever, E 5e5, nu 0.33, gravity, z = 0 fixed,
alpha 0.1, beta 0.001, dt 0.01): plain-solver episodes from rest under seeded
random tractions (0.5e3 .. 2e3 Pa, random free face and direction) are
recorded, the network (which starts as quadratic extrapolation of the
velocity) is trained on them with Adam on the residual loss, and held-out
episodes (other seeds) are stepped with the plain start, the untrained network
and the trained one, CG to relative residual 1e-6 with a cap high enough that
every step converges.  The trained network must need at least 10% fewer CG
iterations than the plain start; the numbers of all three are printed.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_trained_network_cuts_cg_iterations():
    from opentissue_amd import fem, warmstart, workloads
    w = workloads.cantilever()
    s = warmstart.make_solver(w)
    model = fem.Model(seed=0)
    P = fem.config_params()
    train_loads, _ = warmstart.random_loads(w, 8, seed=1)
    n = warmstart.record(s, model, w, train_loads, 50, P)
    assert n == 8 * 50 * s.n_free

    U = fem.config_params(cap=3000)
    test_loads, _ = warmstart.random_loads(w, 2, seed=1001)
    plain = warmstart.run_episodes(s, w, test_loads, 40, U, learned=False)
    init = warmstart.run_episodes(s, w, test_loads, 40, U, learned=True, model=model)
    loss = model.train_residual(s, w.dt, P, steps=250, lr=1e-4, seed=0)
    assert np.all(np.isfinite(loss)) and loss[-50:].mean() < 1.0  # better than the plain start
    learned = warmstart.run_episodes(s, w, test_loads, 40, U, learned=True, model=model)
    print("plain", plain, "\nuntrained", init, "\ntrained", learned)
    assert learned["mean_iterations"] < 0.9 * plain["mean_iterations"]
