"""The C++ drop-in header (<otfem/fem.hpp>) driven by caller code written
against the reference's fem API, compared step by step with the FP64 oracle
(each oracle step starts from the state the C++ program recorded)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_disp_error

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "build", "dropin_c1")


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_cpp_dropin_matches_oracle_per_step(tmp_path, mode):
    import oracle
    from opentissue_amd import mesh
    assert os.path.exists(EXE), "build() compiles tests/cpp/dropin_c1.cpp"
    I, J, K, steps = 4, 8, 3, 5
    out = tmp_path / "x.bin"
    r = subprocess.run([EXE, str(I), str(J), str(K), str(steps), str(mode), str(out)],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    c, t = mesh.generate_blocks(I, J, K, 0.01, 0.01, 0.01)
    N = len(c)
    data = np.fromfile(out, dtype=np.float64)
    states = data[:(steps + 1) * 6 * N].reshape(steps + 1, 2, N, 3)
    nk = int(data[(steps + 1) * 6 * N])
    fixed = (c[:, 2] == 0).astype(np.uint8)
    fext = data[(steps + 1) * 6 * N + 1:].reshape(steps, N, 3)
    o = oracle.Oracle(c, t, fixed)
    if mode == 2:
        o.init(5e5, 0.33, 1000.0, 1e-4, 20.0, 0.05)
    else:
        o.init(5e5, 0.33, 1000.0)
    m = o.mass()
    f = mesh.gravity_forces(np.where(fixed == 1, 0.0, m), fixed) * (50.0 if mode == 2 else 1.0)
    P = oracle.reference_params() if mode == 0 else oracle.config_params()
    if mode == 3:
        P = oracle.config_params(tol=1e-10, cap=5000)
    np.testing.assert_array_equal(states[0, 0], c)
    errs = []
    for k in range(steps):
        # the caller resets m_f_external before every step
        o.set_external_forces(f)
        xs, vs = states[k, 0].copy(), states[k, 1].copy()
        if mode == 3:  # the caller's edits before step k (x only, then v only)
            if k % 2 == 1:
                xs[fixed == 0, 0] += 1e-4 * np.sin(np.arange(N))[fixed == 0]
            if k == 2:
                vs *= 0.5
        o.set_state(xs, vs)
        o.simulate(0.01, True, P)
        xo, _ = o.state()
        errs.append(rel_disp_error(states[k + 1, 0], xo, c))
        # m_f_external read back: the step's plastic forces added to the load
        np.testing.assert_allclose(fext[k], o.external_forces(), rtol=1e-6,
                                   atol=1e-9 * np.abs(f).max())
    print("per-step errors", errs)
    assert max(errs) < 1e-6
    if mode == 2:
        assert np.abs(fext[-1] - f).max() > 1e-3 * np.abs(f).max()  # plasticity acted
    # fem::export_K: an identity entry per fixed dof plus 9 per free-free block
    Kd = o.export_K_dense()
    fr = np.repeat(fixed == 0, 3)
    Kff = Kd[np.ix_(fr, fr)]
    blocks = (np.abs(Kff.reshape(fr.sum() // 3, 3, fr.sum() // 3, 3)).sum((1, 3)) > 0).sum()
    assert nk == 9 * blocks + 3 * int(fixed.sum())


def test_cpp_dropin_timing_mode():
    """dropin_c1 --time: fem::simulate with and without the rotation
    read-back, then otfem_step alone on the same solver."""
    import json
    assert os.path.exists(EXE), "build() compiles tests/cpp/dropin_c1.cpp"
    for rot in (1, 0):
        r = subprocess.run([EXE, "--time", "6", "6", "6", "4", "2", str(rot)],
                           capture_output=True, text=True, timeout=120)
        print(r.stdout, r.stderr)
        assert r.returncode == 0
        d = json.loads(r.stdout.strip().splitlines()[-1])
        assert d["tets"] == 5 * 6 ** 3 and d["rotations"] == rot
        for k in ("dropin_ms_per_step", "direct_ms_per_step", "dropin_kernel_ms_per_step",
                  "direct_kernel_ms_per_step", "dropin_cg_iterations_per_step"):
            assert d[k] > 0, k
