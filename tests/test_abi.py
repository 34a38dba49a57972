"""C ABI checks that need no GPU: the library loads, exports every function
include/otfem/otfem.h declares, validates its inputs and builds the tile
layout on the host."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "otfem", "otfem.h")


@pytest.fixture(scope="module")
def fem():
    lib = os.path.join(ROOT, "opentissue_amd", "libotfem.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-j4", "-C", os.path.join(ROOT, "opentissue_amd", "csrc")], check=True)
    from opentissue_amd import fem as F
    F.load_library()
    return F


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(otfem_[A-Za-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(fem):
    names = declared_functions()
    assert len(names) >= 25
    lib = C.CDLL(fem.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python mirror binds each of them
    assert sorted(fem.SIGNATURES) == names


def test_reference_defaults(fem):
    p = fem.reference_params()
    # fem_simulate.h:104,108-109 and fem_conjugate_gradients.h:43-44
    assert (p.alpha, p.beta, p.cg_min_iterations, p.cg_max_iterations) == (2.0, 0.0, 20, 20)
    assert (p.tol_mode, p.tol, p.tiny, p.rotation_mode, p.warm_start) == (0, 1e-3, 1e-10, 0, 0)
    q = fem.default_params(1)
    assert (q.alpha, q.beta, q.tol, q.cg_max_iterations, q.tol_mode, q.tiny) == (0.1, 0.001, 1e-6, 200, 1, 0.0)
    assert fem.load_library().otfem_default_params(C.byref(p), 7) == fem.INVALID_ARG


def test_create_validates_input(fem):
    from opentissue_amd import mesh
    c, t = mesh.generate_blocks(2, 2, 2, 0.1, 0.1, 0.1)
    bad = t.copy()
    bad[3, 2] = len(c)
    with pytest.raises(fem.OtfemError, match="otfem_create failed: .+"):
        fem.Solver(c, bad)
    bad[3, 2] = -1
    with pytest.raises(fem.OtfemError):
        fem.Solver(c, bad)
    # body offsets must cover the mesh and tets must stay inside their body
    c2, t2, no, to = mesh.pack_bodies([(c, t), (c, t)])
    fem.Solver(c2, t2, body_node_offsets=no, body_tet_offsets=to).close()
    short = no.copy()
    short[-1] -= 1  # offsets must end at num_nodes
    with pytest.raises(fem.OtfemError):
        fem.Solver(c2, t2, body_node_offsets=short, body_tet_offsets=to)
    cross = t2.copy()
    cross[0, 0] = no[1]  # body 0 tet pointing into body 1
    with pytest.raises(fem.OtfemError):
        fem.Solver(c2, cross, body_node_offsets=no, body_tet_offsets=to)


def test_create_reports_why_it_failed(fem):
    from opentissue_amd import mesh
    lib = fem.load_library()
    c, t = mesh.generate_blocks(30, 30, 30, 0.01, 0.01, 0.01)
    # one tile of 29,791 nodes cannot fit a workgroup's LDS
    with pytest.raises(fem.OtfemError, match="LDS budget"):
        fem.Solver(c, t, target_tile_nodes=100000)
    nan = c.copy()
    nan[5, 1] = np.nan
    with pytest.raises(fem.OtfemError, match="non-finite coordinate"):
        fem.Solver(nan, t)
    fem.Solver(c, t).close()
    assert lib.otfem_create_last_error() == b""  # cleared by a success
    assert lib.otfem_create(None, None) == fem.INVALID_ARG
    assert b"null" in lib.otfem_create_last_error()


@pytest.mark.parametrize("dims,tile", [((10, 40, 10), 0), ((6, 7, 5), 64), ((1, 1, 1), 0)])
def test_layout_counts(fem, dims, tile):
    from opentissue_amd import mesh
    c, t = mesh.generate_blocks(*dims, 0.01, 0.01, 0.01)
    s = fem.Solver(c, t, target_tile_nodes=tile)
    info = s.layout_info()
    N, T = len(c), len(t)
    # default for one body this small: min(248, N / 16) tiles
    want = -(-N // tile) if tile else max(1, min(248, N // 16))
    assert info["tiles"] == want
    # every tet has 1..4 copies (one per distinct owning tile of its corners)
    assert T <= info["slots"] <= 4 * T
    assert info["max_tile_nodes"] <= N
    if info["tiles"] == 1:
        assert info["slots"] == T and info["halo"] == 0
    s.close()


def test_batched_layout_never_mixes_bodies(fem):
    from opentissue_amd import mesh
    bodies = [mesh.generate_blocks(a, b, c, 0.01, 0.01, 0.01) for a, b, c in [(1, 1, 1), (3, 2, 4), (2, 2, 2)]]
    c, t, no, to = mesh.pack_bodies(bodies)
    s = fem.Solver(c, t, body_node_offsets=no, body_tet_offsets=to, target_tile_nodes=16)
    info = s.layout_info()
    # tiles per body = ceil(N_b / 16), summed
    assert info["tiles"] == sum(-(-(no[b + 1] - no[b]) // 16) for b in range(3))
    s.close()


def test_layout_independent_of_thread_count():
    """tests/cpp/layout_check.cpp: the host layout is the same for 1, 3 and 8
    threads (batched bodies and one 256-tile body), its slots and halos
    resolve to the mesh's corners, and error precedence is kept."""
    exe = os.path.join(ROOT, "tests", "cpp", "build", "layout_check")
    assert os.path.exists(exe), "build() compiles tests/cpp/layout_check.cpp"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "layout_check ok" in r.stdout


def test_calls_before_init_report_state(fem):
    from opentissue_amd import mesh
    c, t = mesh.generate_blocks(1, 1, 1, 0.1, 0.1, 0.1)
    s = fem.Solver(c, t)
    st = fem.Stats()
    rc = s.lib.otfem_step(s.h, 0.01, 1, None, C.byref(st))
    assert rc == fem.STATE
    m = np.empty(len(c))
    assert s.lib.otfem_get_mass(s.h, m.ctypes.data_as(C.POINTER(C.c_double))) == fem.STATE
    n = C.c_int()
    fem.load_library().otfem_device_count(C.byref(n))
    if n.value == 0:
        assert s.lib.otfem_init(s.h, 5e5, 0.3, 1000.0, 0, 0, 0) == fem.NO_DEVICE
    s.close()


def test_blocks_counts_and_argument_checks(fem):
    """otfem_blocks_count and the argument checks of otfem_blocks_build run
    on the host; without a device the build reports NO_DEVICE."""
    dims = np.array([[3, 4, 5], [1, 1, 1]], dtype=np.int32)
    d = fem.BlocksDesc()
    d.num_bodies = 2
    d.dims = dims.ctypes.data_as(C.POINTER(C.c_int32))
    d.width = d.height = d.depth = 0.01
    N, T = C.c_int64(), C.c_int64()
    assert fem.load_library().otfem_blocks_count(C.byref(d), C.byref(N), C.byref(T)) == fem.OK
    assert N.value == 4 * 5 * 6 + 8 and T.value == 5 * (60 + 1)
    with pytest.raises(fem.OtfemError, match="I, J, K >= 1"):
        fem.build_blocks([(2, 0, 2)], 0.01)
    with pytest.raises(fem.OtfemError, match="edges must be positive"):
        fem.build_blocks([(2, 2, 2)], 0.01, height=-1.0)
    with pytest.raises(ValueError):
        fem.build_blocks([(2, 2, 2)], 0.01, fixed_face="w+")
    with pytest.raises(ValueError):
        fem.build_blocks([(2, 2, 2), (1, 1, 1)], 0.01, load_face="x+", traction=np.ones(4))
    with pytest.raises(fem.OtfemError, match="jitter must be"):
        fem.build_blocks([(2, 2, 2)], 0.01, jitter=0.6)
    n = C.c_int()
    fem.load_library().otfem_device_count(C.byref(n))
    if n.value == 0:
        with pytest.raises(fem.OtfemError) as e:
            fem.build_blocks([(2, 2, 2)], 0.01)
        assert e.value.status == fem.NO_DEVICE


def test_status_strings(fem):
    lib = fem.load_library()
    names = {fem.OK: b"ok", fem.INVALID_ARG: b"invalid argument", fem.STATE: b"invalid state",
             fem.OUT_OF_MEMORY: b"out of host memory"}
    for code, text in names.items():
        assert lib.otfem_status_string(code) == text
    assert lib.otfem_status_string(99) == b"unknown status"


def test_blocks_reject_bad_face_axis(fem):
    dims = np.array([[2, 2, 2]], dtype=np.int32)
    x = np.zeros(3 * 27)
    d = fem.BlocksDesc()
    d.num_bodies = 1
    d.dims = dims.ctypes.data_as(C.POINTER(C.c_int32))
    d.width = d.height = d.depth = 0.01
    d.fixed_axis, d.load_axis = -2, -1
    t = np.zeros(4 * 40, np.int32)
    off = np.zeros(2, np.int64)
    st = fem.load_library().otfem_blocks_build(
        C.byref(d), x.ctypes.data_as(C.POINTER(C.c_double)), t.ctypes.data_as(C.POINTER(C.c_int32)),
        None, None, off.ctypes.data_as(C.POINTER(C.c_int64)), off.ctypes.data_as(C.POINTER(C.c_int64)), None)
    assert st == fem.INVALID_ARG
    assert b"face axis" in fem.load_library().otfem_blocks_last_error()


def test_missing_library_fails_loudly(monkeypatch):
    """Without the built HIP library nothing falls back: every call raises."""
    from opentissue_amd import fem, mesh
    monkeypatch.setattr(fem, "_lib", None)
    monkeypatch.setattr(fem, "LIB_PATH", "/nonexistent/libotfem.so")
    c, t = mesh.generate_blocks(1, 1, 1, 0.01, 0.01, 0.01)
    with pytest.raises(fem.OtfemError, match="missing"):
        fem.Solver(c, t)
    with pytest.raises(fem.OtfemError, match="missing"):
        fem.config_params()
