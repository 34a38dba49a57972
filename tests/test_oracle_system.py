"""Pins of the oracle's assembled system, its stopping rule and its plastic
force against quantities built without the oracle (CPU only).

* The assembled system of fem_dynamics_assembly.h:77-91 (with the stiffness
  damping extension):  A = (1 + alpha dt) M + (dt^2 + beta dt) K',
  b = M v + dt (f_ext - K'(x - x0)) for an unwarped step, checked by a dense
  solve with K from the textbook Voigt element stiffness and M from the
  lattice's closed-form lumped mass.
* The relative-residual stop: at the stop the true residual
  ||b_f - A_ff v_f|| / ||b_f|| of that independent system is <= tol, and one
  iteration earlier it is not.
* The plastic force of fem_add_plasticity_force.h:140-198,
  f_j = Re V B_j^T D e_p: with creep 1/dt the plastic strain equals the total
  strain after one step, so the plastic force cancels the corotated elastic
  force Re Ke (Re^T x - x0) node by node; under a pure stretch its face total
  is sigma_xx times the face area with sigma from Lame's constants.

Each pin is a function that returns its error, so the last test can run it
on oracle variants with planted mistakes (dt^2 -> dt, beta dropped, the
square root dropped from the stop, D0 and D1 swapped, the volume dropped,
Re transposed) and require that each variant fails it.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from opentissue_amd import mesh as M

from test_oracle import _voigt_Ke  # textbook Ke = V B^T D B

H = 0.01
E, NU, RHO = 5e5, 0.33, 1000.0


def _lattice(I=2, J=3, K=2):
    c, t = M.generate_blocks(I, J, K, H, H, H)
    fixed = (np.abs(c[:, 2]) < 1e-12).astype(np.uint8)  # J runs along world z
    return c, t, fixed


def _voigt_K(c0, t, E=E, nu=NU, rot=None):
    """Global rest stiffness from the textbook element matrices, optionally
    warped per element by the rotations rot[e] (R Ke R^T)."""
    n = 3 * len(c0)
    K = np.zeros((n, n))
    for e, tt in enumerate(t):
        Ke = _voigt_Ke(c0[tt], E, nu)
        if rot is not None:
            Rb = np.kron(np.eye(4), rot[e])
            Ke = Rb @ Ke @ Rb.T
        idx = (3 * tt[:, None] + np.arange(3)).ravel()
        K[np.ix_(idx, idx)] += Ke
    return K


def _lumped_mass(c0, t, rho=RHO):
    """m_i = rho h^3 (n_corner / 24 + n_central / 12) on an unjittered
    lattice: corner tets have volume h^3/6, the central tet h^3/3."""
    V = M.tet_volumes(c0, t)
    corner = np.isclose(V, H ** 3 / 6)
    nc = np.bincount(t[corner].ravel(), minlength=len(c0))
    nm = np.bincount(t[~corner].ravel(), minlength=len(c0))
    return rho * H ** 3 * (nc / 24 + nm / 12)


def _system(c0, t, fixed, u, v0, f, dt, alpha, beta):
    K = _voigt_K(c0, t)
    m = np.repeat(_lumped_mass(c0, t), 3)
    A = (1 + alpha * dt) * np.diag(m) + (dt * dt + beta * dt) * K
    b = m * v0.ravel() + dt * (f.ravel() - K @ u.ravel())
    free = np.repeat(fixed == 0, 3)
    return A[np.ix_(free, free)], b[free], free


def _random_state(c, fixed, seed):
    rng = np.random.default_rng(seed)
    mov = (fixed == 0)[:, None]
    u = 2e-3 * H * rng.normal(size=c.shape) * mov
    v0 = 0.02 * rng.normal(size=c.shape) * mov
    f = 1e-3 * rng.normal(size=c.shape) * mov
    return u, v0, f


def pin_assembled_system(oracle_mod, alpha, beta, lib=None):
    """Relative difference between the oracle's converged unwarped step and a
    dense solve of the independently assembled system."""
    c, t, fixed = _lattice()
    u, v0, f = _random_state(c, fixed, 21)
    dt = 0.01
    o = oracle_mod.Oracle(c + u, t, fixed, velocity=v0, model_coords=c, lib=lib)
    o.init(E, NU, RHO)
    o.set_external_forces(f)
    st = o.simulate(dt, False, oracle_mod.config_params(alpha=alpha, beta=beta, tol=1e-13, cap=5000))
    assert st.converged == 1
    Aff, bf, free = _system(c, t, fixed, u, v0, f, dt, alpha, beta)
    vf = np.linalg.solve(Aff, bf)
    _, v = o.state()
    return np.abs(v.ravel()[free] - vf).max() / np.abs(vf).max()


def pin_stop_rule(oracle_mod, lib=None):
    """(true rel-res at the oracle's stop with tol 1e-6, true rel-res when
    the cap is one iteration less, iterations) on the independent system."""
    c, t, fixed = _lattice(3, 4, 3)
    u, v0, f = _random_state(c, fixed, 22)
    dt, alpha, beta = 0.01, 0.1, 0.001
    Aff, bf, free = _system(c, t, fixed, u, v0, f, dt, alpha, beta)

    def run(cap):
        o = oracle_mod.Oracle(c + u, t, fixed, velocity=v0, model_coords=c, lib=lib)
        o.init(E, NU, RHO)
        o.set_external_forces(f)
        st = o.simulate(dt, False, oracle_mod.config_params(alpha=alpha, beta=beta, tol=1e-6, cap=cap))
        _, v = o.state()
        return np.linalg.norm(bf - Aff @ v.ravel()[free]) / np.linalg.norm(bf), st

    res, st = run(5000)
    res_short, _ = run(st.iterations - 1)
    return res, res_short, st.iterations


def pin_plastic_force(oracle_mod, rotated, lib=None):
    """Relative difference between the plastic force after one step with
    creep 1/dt (the oracle adds it into f_external, as the reference does)
    and Re Ke (Re^T x - x0) from the textbook Ke and the step's rotations;
    under a pure stretch also the x = max face total against sigma_xx A."""
    c, t, fixed = _lattice(3, 3, 3)
    eps = np.array([2e-3, -1e-3, 5e-4])
    X = c * (1 + eps)
    if rotated:
        rng = np.random.default_rng(23)
        Q = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        Q *= np.sign(np.linalg.det(Q))
        X = (X + 1e-3 * H * rng.normal(size=c.shape)) @ Q.T + 0.01
    dt = 0.01
    o = oracle_mod.Oracle(X, t, model_coords=c, lib=lib)
    o.init(E, NU, RHO, 0.0, 1.0 / dt, 1.0)  # yield 0, creep 1/dt, max far above the strain
    o.simulate(dt, True, oracle_mod.config_params(tol=1e-12, cap=2000))
    fp = o.external_forces()
    R = o.rotations()
    fe = np.zeros_like(c)
    for e, tt in enumerate(t):
        Ke = _voigt_Ke(c[tt], E, NU)
        ue = (X[tt] @ R[e] - c[tt]).ravel()  # Re^T x - x0 per corner
        fe[tt] += (np.kron(np.eye(4), R[e]) @ (Ke @ ue)).reshape(4, 3)
    err = np.abs(fp - fe).max() / np.abs(fe).max()
    if rotated:
        return err, 0.0
    lam, mu = E * NU / ((1 + NU) * (1 - 2 * NU)), E / (2 * (1 + NU))
    sxx = lam * eps.sum() + 2 * mu * eps[0]
    hi = c.max(0)
    face = np.isclose(c[:, 0], hi[0])
    area = (hi[1] - c[:, 1].min()) * (hi[2] - c[:, 2].min())
    return err, abs(fp[face, 0].sum() - sxx * area) / (sxx * area)


@pytest.mark.parametrize("alpha,beta", [(2.0, 0.0), (0.1, 0.001)])
def test_assembled_system_equals_dense_solve(oracle_mod, alpha, beta):
    assert pin_assembled_system(oracle_mod, alpha, beta) < 1e-9


def test_relative_residual_stop(oracle_mod):
    res, res_short, its = pin_stop_rule(oracle_mod)
    print("stop at", its, "iterations: true rel-res", res, "one iteration less", res_short)
    assert its > 10
    assert res <= 1e-6 * (1 + 1e-6)
    assert res_short > 1e-6


@pytest.mark.parametrize("rotated", [False, True])
def test_plastic_force_cancels_elastic_force(oracle_mod, rotated):
    err, face = pin_plastic_force(oracle_mod, rotated)
    assert err < 1e-9
    assert face < 1e-9


# Planted mistakes: (description, original text, replacement).  Each must
# make at least one pin above fail.
MUTATIONS = [
    ("dt^2 -> dt in A", "const double kscale = dt * dt + beta * dt;", "const double kscale = dt + beta * dt;"),
    ("beta dropped", "const double kscale = dt * dt + beta * dt;", "const double kscale = dt * dt;"),
    ("mass damping without dt", "double tmp = m_i + dt * c_i;", "double tmp = m_i + c_i;"),
    ("sqrt dropped from the stop", "std::sqrt(d1) <= rel_stop", "d1 <= P.tol * bb"),
    ("D0 and D1 swapped in the plastic force", "double E0 = T.D(0), E1 = T.D(1), E2 = T.D(2);",
     "double E0 = T.D(1), E1 = T.D(0), E2 = T.D(2);"),
    ("volume dropped from the plastic force", "f *= T.V;", "f *= 1.0;"),
    ("Re transposed in the plastic force", "m.nodes[T.n[j]].f_external += T.Re * f;",
     "m.nodes[T.n[j]].f_external += trans(T.Re) * f;"),
]


def _fails(oracle_mod, lib):
    for ab in [(2.0, 0.0), (0.1, 0.001)]:
        if not pin_assembled_system(oracle_mod, *ab, lib=lib) < 1e-9:
            return True
    res, res_short, _ = pin_stop_rule(oracle_mod, lib=lib)
    if not (res <= 1e-6 * (1 + 1e-6) and res_short > 1e-6):
        return True
    for rotated in (False, True):
        err, face = pin_plastic_force(oracle_mod, rotated, lib=lib)
        if not (err < 1e-9 and face < 1e-9):
            return True
    return False


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ to build oracle variants")
def test_pins_catch_planted_mistakes(oracle_mod, tmp_path):
    """Builds one oracle variant per planted mistake (g++, CPU) and checks
    that the pins above reject every one of them."""
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    src = open(os.path.join(here, "fem_oracle.cpp")).read()
    missed = []
    for k, (what, old, new) in enumerate(MUTATIONS):
        assert src.count(old) == 1, what
        path = tmp_path / f"fem_oracle_m{k}.cpp"
        path.write_text(src.replace(old, new))
        so = tmp_path / f"liboracle_m{k}.so"
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-shared",
                        "-o", str(so), str(path)], check=True, timeout=600)
        lib = oracle_mod.load_library(str(so))
        if not _fails(oracle_mod, lib):
            missed.append(what)
    assert not missed, f"pins did not catch: {missed}"
