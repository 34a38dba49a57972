#!/usr/bin/env python3
"""Throughput of the corotated tetrahedral FEM step on one MI355X.

Default workload: the combined set -- 2,048 uneven jittered lattice bodies
(dims uniform in [4, 32] cubes per axis, jitter +/-0.25 h with resampling,
per-body E log-uniform [1e4, 1e6], nu uniform [0.25, 0.45], one face fixed
and a random traction of 0.5e3..2e3 Pa on the opposite face), dt 0.01,
Rayleigh alpha 0.1 / beta 0.001, CG run per body until
|b - A v| <= 1e-6 |b| (a safety cap of 5,000 iterations), all bodies
stepped together (CSR of bodies, per-body CG state), tiles chosen by the
layout autotuner.  A step is one full implicit step (rotations + elastic
forces + right-hand side + CG start + CG + position update) of every body,
inputs resident in HBM.

Legs, each from rest: --warmup untimed steps, then exactly --steps timed
steps (steps W+1 .. W+K of the episode):
  learned   CG started from the learned start: the subspace start over the
            last --subspace velocity differences with the network's
            prediction as one more direction (the network is trained first,
            untimed, on 2,000 plain-solver steps of the 10x10x40 cantilever
            under seeded random tractions, Adam lr 1e-3, 65,536-node
            minibatches, 2,000 updates)            -> `value`
  plain     CG started from v^n, the reference's start
  ablations the subspace start without the network, the network without
            the subspace start
`episode` times the workload's whole 100-step episode from rest for plain
and learned.  `roofline` prices the CG kernel (one k_cg_iter launch per CG
iteration, each body at its own iteration count: 100 B/tet + 195 B/node,
FP64 v and r, plus 36 B/node for the position update) against the MI355X
HBM peak.  `cpu_baseline` times the FP64 CPU restatement of the reference
step (one thread) on a bounded sample of the same workload;
`cpu_baseline_c1` is the 10x10x40 cantilever's 100-step run in both modes.
`secondary` repeats plain and learned on the 64^3 cube.

    python bench.py [--steps 20] [--warmup 5] [--config combined|bodies|cube64|cantilever]
    python bench.py --copy-probe     # plain device copy of known size, for
                                     # scaling rocprofv3 FETCH/WRITE counters
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_TET_ITER = 100
BYTES_PER_NODE_ITER = 195
BYTES_PER_NODE_POSUPDATE = 36
MFMA_BF16_PEAK_TFLOPS = 2500.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def build_workload(name):
    from opentissue_amd import workloads
    if name == "cube64":
        return workloads.cube64()
    if name == "cantilever":
        return workloads.cantilever()
    if name == "bodies":
        return workloads.bodies(512, 4, 24)
    if name == "combined":
        return workloads.combined(2048, 4, 32)
    raise SystemExit(f"unknown config {name}")


def copy_probe():
    import torch
    n = 1 << 28  # 2^28 float32 = 1 GiB
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n,
                      "launches": 5}))


def stream_rate():
    """A streaming reference on this GPU: c = a + b over 2 GiB arrays (two
    reads and one write per element), GB/s of bytes moved: the rate a
    bandwidth-bound kernel can expect here.  Measured in a child process
    (torch brings its own HIP runtime, which does not start in a process
    where the library's runtime already holds the GPU); None if it fails."""
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--stream-probe"],
                       capture_output=True, text=True, timeout=300)
    try:
        return float(json.loads(r.stdout.strip().splitlines()[-1])["stream_gbs"])
    except (ValueError, IndexError, KeyError):
        log(f"streaming reference failed: {r.stderr.strip()[-300:]}")
        return None


def stream_probe():
    """The child of stream_rate: HIP events over 10 launches of c = a + b."""
    import torch
    n = 1 << 29  # 2 GiB of float32 per array
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.ones_like(a)
    c = torch.empty_like(a)
    torch.add(a, b, out=c)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        torch.add(a, b, out=c)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) / 1e3 / 10
    print(json.dumps({"stream_gbs": 3 * 4 * n / sec / 1e9}))


# ---------------------------------------------------------------------------
# CPU baselines (the FP64 oracle, one thread)

def run_cpu_baseline(w, params_kw):
    """One implicit step of the same workload from rest with the FP64 oracle."""
    import oracle
    o = oracle.Oracle(w.coords, w.tets, w.fixed)
    o.init_materials(w.young, w.poisson, w.density)
    o.set_external_forces(w.forces)
    P = oracle.config_params(**params_kw)
    t0 = time.perf_counter()
    st = o.simulate(w.dt, True, P)
    t = time.perf_counter() - t0
    x, _ = o.state()
    return dict(seconds=t, iterations=st.iterations, x=x)


def run_cpu_baseline_bodies(w, params_kw, seconds_budget, xg):
    """Bodies stepped one by one (each its own CG, as on the device) with the
    FP64 oracle from rest until the time budget is spent; also the first-step
    displacement error of the device against those bodies."""
    import oracle
    P = oracle.config_params(**params_kw)
    no, to = w.body_node_offsets, w.body_tet_offsets
    secs, tets, its, bodies = 0.0, 0, 0, 0
    err = {"converged": 0.0, "capped": 0.0}
    n_capped = 0
    for b in range(w.num_bodies):
        ns, ts = slice(no[b], no[b + 1]), slice(to[b], to[b + 1])
        c = w.coords[ns]
        o = oracle.Oracle(c, w.tets[ts] - no[b], w.fixed[ns])
        o.init_materials(w.young[ts], w.poisson[ts], w.density[ts])
        o.set_external_forces(w.forces[ns])
        t0 = time.perf_counter()
        st = o.simulate(w.dt, True, P)
        secs += time.perf_counter() - t0
        x, _ = o.state()
        den = np.linalg.norm(x - c, axis=1).max()
        key = "converged" if st.converged else "capped"
        n_capped += 0 if st.converged else 1
        if den > 0:
            err[key] = max(err[key], np.linalg.norm(xg[ns] - x, axis=1).max() / den)
        tets += int(to[b + 1] - to[b])
        its += st.iterations
        bodies += 1
        if secs >= seconds_budget:
            break
    return dict(seconds=secs, tets=tets, bodies=bodies, iterations=its, err=err,
                capped=n_capped)


def cpu_baseline_c1():
    """The 10x10x40 cantilever (20,000 tets, gravity, z = 0 fixed), 100 steps
    from rest with the FP64 oracle on one thread, in the relative-residual
    mode (alpha 0.1, beta 0.001, rel-res 1e-6, cap 200) and in the reference
    mode (alpha 2, beta 0, 20 CG iterations, the clamps)."""
    import oracle
    from opentissue_amd import workloads
    w = workloads.cantilever()
    out = {}
    for mode, P in (("config", oracle.config_params(cap=200)),
                    ("reference", oracle.reference_params())):
        o = oracle.Oracle(w.coords, w.tets, w.fixed)
        o.init_materials(w.young, w.poisson, w.density)
        o.set_external_forces(w.forces)
        its = 0
        t0 = time.perf_counter()
        for _ in range(100):
            its += o.simulate(w.dt, True, P).iterations
        t = time.perf_counter() - t0
        out[mode] = {"tet_steps_per_s": w.num_tets * 100 / t, "seconds": t,
                     "cg_iterations_per_step": its / 100}
    model, ncpu = cpu_info()
    out.update({"unit": "tet-steps/s", "cores": 1, "kind": "port",
                "sample": f"{w.name} ({w.num_tets} tets), 100 steps from rest, FP64 ordered-map "
                          f"restatement of fem::simulate on {model} ({ncpu} host cpus, 1 used)"})
    return out


# ---------------------------------------------------------------------------
# learned start: training and the cantilever A/B

def train_networks():
    """Networks for the learned start, trained on 2,000 plain-solver steps
    (20 episodes of 100) of the 10x10x40 cantilever under seeded random
    tractions with Adam on the residual loss: the configuration's
    hyperparameters (lr 1e-3, minibatches of whole recorded steps closest to
    65,536 nodes, 2,000 updates) and the tuned ones (lr 1e-4, 250 updates).
    Returns (models, solver, workload, info)."""
    from opentissue_amd import fem, warmstart, workloads
    w = workloads.cantilever()
    s = warmstart.make_solver(w)
    P = fem.config_params()
    rec = fem.Model(seed=0)
    t0 = time.perf_counter()
    loads, _ = warmstart.random_loads(w, 20, seed=1)
    n = warmstart.record(s, rec, w, loads, 100, P)
    t1 = time.perf_counter()
    models, info = {}, {"samples": int(n), "record_seconds": t1 - t0,
                        "nodes_per_batch": int(s.n_free * max(1, round(65536 / s.n_free)))}
    theta0 = rec.params()
    for name, lr, steps in (("config", 1e-3, 2000), ("tuned", 1e-4, 250)):
        if name == "config":
            m = rec
        else:
            # the residual loss reads the steps `s` recorded into the model's
            # own store: the second network records the same episodes
            m = fem.Model(seed=0)
            warmstart.record(s, m, w, loads, 100, P)
        t2 = time.perf_counter()
        loss = m.train_residual(s, w.dt, P, steps=steps, lr=lr, seed=0)
        info[name] = {"lr": lr, "adam_steps": steps, "train_seconds": time.perf_counter() - t2,
                      "final_loss": float(np.mean(loss[-min(50, steps):]))}
        models[name] = m
    untrained = fem.Model(seed=0)
    untrained.set_params(theta0)
    models["untrained"] = untrained
    return models, s, w, info


def cantilever_ab(models, s, w, subspace):
    """Held-out cantilever episodes (seed 1001, 3 x 100 steps from rest), CG
    to rel-res 1e-6 with cap 3000 (every step converges): CG iterations,
    throughput and the start residual |r0| / |b| per start."""
    from opentissue_amd import fem
    from opentissue_amd import warmstart
    loads, _ = warmstart.random_loads(w, 3, seed=1001)
    variants = [("plain", None, 0), ("network_untrained", "untrained", 0),
                ("network_tuned", "tuned", 0), ("network_config", "config", 0),
                ("subspace", None, subspace), ("learned", "config", subspace)]
    out = {"workload": f"{w.name} ({w.num_tets} tets), 3 held-out episodes x 100 steps from rest, "
                       "random traction 0.5e3..2e3 Pa, CG rel-res 1e-6 cap 3000"}
    for name, mname, sub in variants:
        P = fem.config_params(cap=3000, subspace=sub)
        model = models[mname] if mname else None
        P.warm_start = fem.WARM_LEARNED if model is not None else fem.WARM_PREVIOUS_VELOCITY
        s.set_warmstart_model(model)
        its, rel0, secs = [], [], 0.0
        for f in loads:
            s.set_state(w.coords, np.zeros_like(w.coords))
            s.set_external_forces(f)
            t0 = time.perf_counter()
            for _ in range(100):
                st = s.simulate(w.dt, True, P)
                its.append(st.iterations)
                rel0.append(np.sqrt(st.rr0 / st.bb) if st.bb > 0 else 0.0)
            secs += time.perf_counter() - t0
        its, rel0 = np.array(its, float), np.array(rel0)
        out[name] = {"cg_iterations_per_step": float(its.mean()),
                     "cg_iterations_steps_11_100": float(its.reshape(3, 100)[:, 10:].mean()),
                     "tet_steps_per_s": w.num_tets * 300 / secs,
                     "start_rel_residual_median": float(np.median(rel0)),
                     "start_rel_residual_p90": float(np.percentile(rel0, 90))}
    s.set_warmstart_model(None)
    p = out["plain"]["cg_iterations_per_step"]
    for name, _, _ in variants[1:]:
        out[name]["iterations_vs_plain"] = p / out[name]["cg_iterations_per_step"]
    return out


# ---------------------------------------------------------------------------
# device legs

def cg_bytes(w, iters, steps):
    """Algorithmic bytes of the CG launches: each body priced at its own
    iteration count, plus the position update."""
    if w.num_bodies == 1:
        b = float(iters.sum()) * (BYTES_PER_TET_ITER * w.num_tets + BYTES_PER_NODE_ITER * w.num_nodes)
    else:
        bt = np.diff(w.body_tet_offsets)
        bn = np.diff(w.body_node_offsets)
        b = float(np.sum(iters * (BYTES_PER_TET_ITER * bt + BYTES_PER_NODE_ITER * bn)))
    return b + steps * BYTES_PER_NODE_POSUPDATE * w.num_nodes


def run_leg(s, w, P, warmup, steps, model=None):
    """Steps from rest: `warmup` untimed, then `steps` timed (device-resident,
    asynchronous, one synchronisation at the end)."""
    s.set_warmstart_model(model)
    s.set_state(w.coords, np.zeros_like(w.coords))
    s.set_external_forces(w.forces)
    for _ in range(warmup):
        s.step_async(w.dt, True, P)
    s.synchronize()
    s.iteration_totals(reset=True)
    s.unconverged_totals(reset=True)
    s.set_timing(True)
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step_async(w.dt, True, P)
    st = s.synchronize()
    wall = time.perf_counter() - t0
    td = s.timing_detail()
    iters = s.iteration_totals().astype(np.float64)
    unconv = int(s.unconverged_totals().sum())
    s.set_timing(False)
    s.set_warmstart_model(None)
    T = w.num_tets
    bt = np.diff(w.body_tet_offsets) if w.num_bodies > 1 else np.array([T])
    nsteps = max(1, td["steps"])
    cgb = cg_bytes(w, iters, steps)
    return {
        "value": T * steps / wall,
        "wall_s": wall,
        "ms_per_step": 1e3 * wall / steps,
        "cg_iterations_per_step": float(iters.sum()) / steps / w.num_bodies,
        "cg_iterations_per_step_tet_weighted": float(np.sum(iters * bt) / bt.sum()) / steps,
        "cg_iterations_per_step_max_body": float(iters.max()) / steps,
        "cg_body_steps_unconverged": unconv,
        "flags": int(st.flags),
        "kernel_ms_per_step": {k: td[k] / nsteps for k in ("network", "element", "subspace", "cg")},
        "cg_algorithmic_bytes_per_step": cgb / steps,
        "cg_hbm_gbs": cgb / (td["cg"] / 1e3) / 1e9 if td["cg"] > 0 else 0.0,
        "_iters": iters,
    }


def setup_solver(w, P, tile_nodes, cg_block=0):
    from opentissue_amd import fem
    t0 = time.perf_counter()
    tuned_block = 0
    if tile_nodes < 0:
        tile_nodes, tuned_block, _ = fem.autotune_layout(
            w.coords, w.tets, w.fixed, w.young, w.poisson, w.density, w.forces, w.dt, P, steps=2,
            body_node_offsets=w.body_node_offsets, body_tet_offsets=w.body_tet_offsets)
        log(f"layout autotuner: target_tile_nodes {tile_nodes}, CG block {tuned_block} "
            f"({time.perf_counter() - t0:.1f} s)")
    t1 = time.perf_counter()
    s = fem.Solver(w.coords, w.tets, w.fixed, body_node_offsets=w.body_node_offsets,
                   body_tet_offsets=w.body_tet_offsets, target_tile_nodes=tile_nodes)
    create_s = time.perf_counter() - t1
    log(f"otfem_create (tile layout) {create_s:.2f} s")
    t1 = time.perf_counter()
    s.init_materials(w.young, w.poisson, w.density)
    init_s = time.perf_counter() - t1
    log(f"otfem_init_materials (element data, masses, bounds on the host) {init_s:.2f} s")
    s.set_external_forces(w.forces)
    block = cg_block or tuned_block or s.autotune(w.dt, True, P, steps=2)
    s.set_kernel_config(cg_block=block)
    s.setup_seconds = {"otfem_create": create_s, "otfem_init_materials": init_s}
    return s, tile_nodes, block, tile_nodes < 0


def measure(w, args, models, tile_nodes, legs=("learned", "plain", "subspace", "network")):
    from opentissue_amd import fem
    params_kw = dict(alpha=0.1, beta=0.001, tol=1e-6, cap=args.cap)
    P = fem.config_params(**params_kw)
    # the layout autotuner skips tilings whose tiles cannot hold the
    # subspace start's directions
    s, tile_nodes, block, _ = setup_solver(w, fem.config_params(**params_kw, subspace=args.subspace),
                                           tile_nodes, args.cg_block)
    if args.deterministic:
        s.set_deterministic(True)
    layout = s.layout_info()
    out = {"tile_nodes": tile_nodes, "cg_block": block, "layout": layout,
           "setup_seconds": s.setup_seconds}
    for leg in legs:
        net = models.get("config") if leg in ("learned", "network") else None
        sub = args.subspace if leg in ("learned", "subspace") else 0
        PL = fem.config_params(**params_kw, subspace=sub)
        PL.warm_start = fem.WARM_LEARNED if net is not None else fem.WARM_PREVIOUS_VELOCITY
        r = run_leg(s, w, PL, args.warmup, args.steps, model=net)
        out[leg] = r
        log(f"{w.name} {leg}: {r['value']:.4g} tet-steps/s, {r['ms_per_step']:.2f} ms/step, "
            f"CG iterations/step {r['cg_iterations_per_step']:.1f} (tet-weighted "
            f"{r['cg_iterations_per_step_tet_weighted']:.1f}, slowest body "
            f"{r['cg_iterations_per_step_max_body']:.1f}), kernels {r['kernel_ms_per_step']}")
    return s, P, params_kw, out


def roofline(w, leg, profile_key):
    """HBM roofline of k_cg_iter from a leg's own HIP events and iteration
    counts; counter figures from the committed rocprofv3 summaries of this
    workload, when there are any (marked as such)."""
    traffic = profile_avg_us = traffic_build = profile_build = None
    tpath = os.path.join(ROOT, "profiles", f"k_cg_iter_{profile_key}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic, traffic_build = tj.get("hbm_bytes_per_launch"), tj.get("build")
    kpath = os.path.join(ROOT, "profiles", f"k_cg_iter_{profile_key}_kernel_trace.json")
    if os.path.exists(kpath):
        with open(kpath) as f:
            kj = json.load(f)
        profile_avg_us, profile_build = kj.get("avg_us_per_working_launch"), kj.get("build")
    achieved = leg["cg_hbm_gbs"]
    out = {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "frac_note": "algorithmic bytes (100 B/tet + 195 B/node per CG iteration of each body, "
                     "36 B/node position update) over the CG kernels' HIP-event time",
        "traffic": traffic,
        "kernel": "k_cg_iter (one launch per CG iteration) + k_finish",
        "algorithmic_bytes_per_step": leg["cg_algorithmic_bytes_per_step"],
        "cg_ms_per_step": leg["kernel_ms_per_step"]["cg"],
    }
    if w.num_bodies == 1 and leg["_iters"].sum() > 0:
        n_it = float(leg["_iters"].sum())
        steps = leg["wall_s"] / leg["ms_per_step"] * 1e3
        out["algorithmic_bytes_per_launch"] = BYTES_PER_TET_ITER * w.num_tets + BYTES_PER_NODE_ITER * w.num_nodes
        out["avg_launch_us"] = 1e3 * leg["kernel_ms_per_step"]["cg"] * steps / n_it
        out["avg_launch_note"] = ("HIP events over the timed CG work divided by the CG iterations: one "
                                  "k_cg_iter launch plus its share of k_finish, the last chunk's "
                                  "early-exit launches and launch gaps")
    if traffic:
        out["traffic_note"] = ("HBM bytes per working k_cg_iter launch from the committed rocprofv3 "
                               "FETCH_SIZE / WRITE_SIZE summary (profiles/), scaled by a copy of known "
                               f"size; not measured in this run; profiled build: {traffic_build}")
        if "avg_launch_us" in out:
            out["counted_frac"] = traffic / (out["avg_launch_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS
        elif profile_avg_us:
            # batched: bytes and duration of a full-grid launch, both from the profiles
            out["counted_frac"] = traffic / (profile_avg_us * 1e-6) / 1e9 / HBM_PEAK_GBS
            out["counted_note"] = ("full-grid k_cg_iter launches: counted HBM bytes over the kernel "
                                   "trace's mean duration, both from the committed profiles")
    if profile_avg_us:
        out["profile_avg_launch_us"] = profile_avg_us
        out["profile_note"] = ("committed rocprofv3 kernel trace (profiles/), not measured in this run; "
                               f"profiled build: {profile_build}")
    return out


def public(leg):
    return {k: v for k, v in leg.items() if not k.startswith("_")}


def episode(s, w, args, models):
    """The workload's whole episode (w.steps steps from rest), plain and learned."""
    from opentissue_amd import fem
    out = {}
    for leg in ("plain", "learned"):
        net = models.get("config") if leg == "learned" else None
        P = fem.config_params(cap=args.cap, subspace=args.subspace if leg == "learned" else 0)
        P.warm_start = fem.WARM_LEARNED if net is not None else fem.WARM_PREVIOUS_VELOCITY
        r = run_leg(s, w, P, 0, w.steps, model=net)
        out[leg] = {"tet_steps_per_s": r["value"], "ms_per_step": r["ms_per_step"],
                    "cg_iterations_per_step_tet_weighted": r["cg_iterations_per_step_tet_weighted"],
                    "cg_body_steps_unconverged": r["cg_body_steps_unconverged"]}
        log(f"{w.name} episode {leg}: {r['value']:.4g} tet-steps/s over {w.steps} steps")
    out["learned_vs_plain"] = out["learned"]["tet_steps_per_s"] / out["plain"]["tet_steps_per_s"]
    out["steps"] = w.steps
    return out


def _wquantile(x, wts, q):
    o = np.argsort(x)
    c = np.cumsum(wts[o])
    return float(x[o][min(len(x) - 1, np.searchsorted(c, q * c[-1]))])


def start_residuals(s, w, args, models):
    """The timed steps of the plain and learned legs again, synchronously:
    per body and step the start residual |b - A v_start| / |b|, the final
    one and the CG iterations, tet-weighted over bodies and steps.  CG's
    iterations per decade of residual reduction belong to the operator; a
    start only removes the decades between the plain start's residual and
    its own, so at equal iterations per decade the start can cut CG
    iterations by at most decades(plain) / decades(learned)."""
    from opentissue_amd import fem
    bt = (np.diff(w.body_tet_offsets) if w.num_bodies > 1 else np.array([w.num_tets])).astype(float)
    out = {"steps": f"{args.warmup + 1}..{args.warmup + args.steps} from rest"}
    for leg in ("plain", "learned"):
        net = models.get("config") if leg == "learned" else None
        P = fem.config_params(cap=args.cap, subspace=args.subspace if leg == "learned" else 0)
        P.warm_start = fem.WARM_LEARNED if net is not None else fem.WARM_PREVIOUS_VELOCITY
        s.set_warmstart_model(net)
        s.set_state(w.coords, np.zeros_like(w.coords))
        s.set_external_forces(w.forces)
        r0, rf, it, wt = [], [], [], []
        for k in range(args.warmup + args.steps):
            s.simulate(w.dt, True, P)
            if k >= args.warmup:
                i, rel = s.body_stats()
                r0.append(s.start_residuals())
                rf.append(rel)
                it.append(i.astype(float))
                wt.append(bt)
        s.set_warmstart_model(None)
        r0, rf, it, wt = (np.concatenate(a) for a in (r0, rf, it, wt))
        ok = (r0 > 0) & (rf > 0)
        dec = np.where(ok, np.log10(np.maximum(r0, 1e-300) / np.maximum(rf, 1e-300)), 0.0)
        out[leg] = {
            "start_rel_residual_p10": _wquantile(r0, wt, 0.1),
            "start_rel_residual_median": _wquantile(r0, wt, 0.5),
            "start_rel_residual_p90": _wquantile(r0, wt, 0.9),
            "final_rel_residual_median": _wquantile(rf, wt, 0.5),
            "decades_per_step": float(np.sum(wt * dec) / wt.sum()),
            "cg_iterations_per_step": float(np.sum(wt * it) / wt.sum()),
            "iterations_per_decade": float(np.sum(wt * it) / max(np.sum(wt * dec), 1e-300)),
        }
        log(f"{w.name} start residuals {leg}: median {out[leg]['start_rel_residual_median']:.3g}, "
            f"{out[leg]['decades_per_step']:.2f} decades and "
            f"{out[leg]['cg_iterations_per_step']:.1f} iterations per step (tet-weighted)")
    p, l = out["plain"], out["learned"]
    out["iterations_ratio"] = p["cg_iterations_per_step"] / l["cg_iterations_per_step"]
    out["decades_ratio"] = p["decades_per_step"] / l["decades_per_step"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="combined")
    ap.add_argument("--subspace", type=int, default=10,
                    help="velocity differences in the learned (subspace) CG start")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="time budget of the multi-body CPU sample")
    ap.add_argument("--tile-nodes", type=int, default=-1,
                    help="owned nodes per tile (0: the default tiling; -1: the layout autotuner "
                         "picks the tile target and the CG block size)")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--stream-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-learned", action="store_true", help="plain leg only")
    ap.add_argument("--legs", default="learned,plain,subspace,network",
                    help="legs to time (learned: network + subspace start; subspace: without the "
                         "network; network: without the subspace start); the first is `value`")
    ap.add_argument("--no-episode", action="store_true", help="skip the 100-step episodes")
    ap.add_argument("--no-secondary", action="store_true", help="skip the 64^3-cube legs")
    ap.add_argument("--cap", type=int, default=5000, help="CG iteration cap (safety net)")
    ap.add_argument("--deterministic", action="store_true",
                    help="per-node sums in slot order (same bits on every run)")
    ap.add_argument("--cg-block", type=int, default=0, choices=(0, 256, 512, 1024),
                    help="threads per tile in the CG kernels (0: autotune)")
    ap.add_argument("--profile", action="store_true",
                    help="one configuration for a profiler: the first leg only, the given layout "
                         "(--tile-nodes, default tiling when -1) and CG block (512 when 0), no "
                         "autotuning, no parameter changes between steps, no other legs")
    args = ap.parse_args()
    if args.profile:
        args.no_cpu = args.no_episode = args.no_secondary = True
        args.tile_nodes = max(args.tile_nodes, 0)
        args.cg_block = args.cg_block or 512

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU "
              "sharding of this path is a design note, not a measured configuration.",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.stream_probe:
        stream_probe()
        return

    legs = ("plain",) if args.no_learned else tuple(args.legs.split(","))
    need_net = "learned" in legs or "network" in legs
    models, ab, train_info = {}, None, None
    if need_net:
        log("learned start: recording 2,000 cantilever steps and training (not timed)")
        models, cs, cw, train_info = train_networks()
        ab = cantilever_ab(models, cs, cw, args.subspace)
        log("cantilever A/B: " + ", ".join(f"{k} {v['cg_iterations_per_step']:.1f} it"
                                            for k, v in ab.items() if isinstance(v, dict)))
        cs.close()

    t_setup = time.perf_counter()
    w = build_workload(args.config)
    log(f"workload {w.name}: {w.num_bodies} bodies, {w.num_tets} tets, {w.num_nodes} nodes "
        f"(built in {time.perf_counter() - t_setup:.2f} s)")
    if args.profile:
        legs = legs[:1]
    elif "plain" not in legs:
        legs = legs + ("plain",)
    s, P, params_kw, m = measure(w, args, models, args.tile_nodes, legs)
    head = m[legs[0]]
    T, N = w.num_tets, w.num_nodes

    result = {
        "metric": "tet-steps/s at CG rel-res 1e-6, learned vs plain; CG HBM GB/s % of peak",
        "value": head["value"],
        "unit": "tet-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {
            "workload": f"{w.name}: {w.num_bodies} bodies, {T} tets, {N} nodes, dt {w.dt}, "
                        f"CG rel-res 1e-6 per body (cap {args.cap}), alpha 0.1 beta 0.001, "
                        + {"plain": "plain start v^n",
                           "learned": f"learned start (network + subspace of {args.subspace} velocity "
                                      "differences) shared by all bodies",
                           "subspace": f"subspace start over {args.subspace} velocity differences",
                           "network": "network start without the subspace"}[legs[0]]
                        + f", steps {args.warmup + 1}.."
                        f"{args.warmup + args.steps} from rest",
            "tets": T, "nodes": N, "bodies": w.num_bodies,
            "tiles": m["layout"]["tiles"], "slots": m["layout"]["slots"], "cg_block": m["cg_block"],
            "tile_nodes": m["tile_nodes"], "layout_autotuned": args.tile_nodes < 0,
            "deterministic": bool(args.deterministic),
            "setup_seconds": m["setup_seconds"],
        },
        "leg": legs[0],
        "learned_vs_plain": (m["learned"]["value"] / m["plain"]["value"]
                             if "learned" in m and "plain" in m else None),
        "plain": public(m["plain"]) if "plain" in m else None,
        "roofline": roofline(w, head, args.config),
        "roofline_plain": roofline(w, m["plain"], args.config) if "plain" in m else None,
        "cpu_baseline": None,
    }
    for leg in legs:
        if leg != "plain":
            result[leg if leg == "learned" else f"ablation_{leg}"] = public(m[leg])
    if "network" in m or "learned" in m:
        ws_ms = m["network" if "network" in m else "learned"]["kernel_ms_per_step"]["network"]
        net_flops = 2.0 * 34688 * N
        net_tflops = net_flops / (ws_ms / 1e3) / 1e12 if ws_ms > 0 else 0.0
        result["network_roofline"] = {
            "bound": "mfma", "kernel": "k_ws_predict", "achieved": net_tflops,
            "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": net_tflops / MFMA_BF16_PEAK_TFLOPS,
            "algorithmic_flops_per_launch": net_flops, "avg_launch_ms": ws_ms,
            "note": "nominal BF16 flops (2 x 34,688 MAC per node); the kernel issues 2.2x as many "
                    "(inputs and activations as BF16 hi + lo pairs, 3-output layer padded to 32 rows)"}
        result["training"] = train_info
        result["cantilever_ab"] = ab
    sr = None if args.profile else stream_rate()
    if sr:
        result["stream_reference"] = {
            "gbs": sr, "frac_of_peak": sr / HBM_PEAK_GBS,
            "kernel": "torch add c = a + b, 2 GiB float32 arrays (2 reads + 1 write per element)",
            "note": "measured in this run; the counted HBM rate of k_cg_iter (profiles/) sits "
                    "against this, its algorithmic rate against the 8 TB/s peak"}
        log(f"streaming reference (c = a + b, 2 GiB arrays): {sr:.0f} GB/s")
    if not args.no_episode and "learned" in legs:
        result["episode"] = episode(s, w, args, models)
        result["start_residuals"] = start_residuals(s, w, args, models)

    if not args.no_cpu:
        if w.num_bodies > 1:
            # bounded sample: bodies of the same workload, one step from rest each
            s.set_state(w.coords, np.zeros_like(w.coords))
            s.set_external_forces(w.forces)
            s.simulate(w.dt, True, P)
            xg, _ = s.state()
            log(f"cpu baseline: bodies of the same workload, one step each, FP64 oracle (1 thread), "
                f"{args.cpu_seconds:.0f} s budget")
            cb = run_cpu_baseline_bodies(w, params_kw, args.cpu_seconds, xg)
            model, ncpu = cpu_info()
            result["cpu_baseline"] = {
                "value": cb["tets"] / cb["seconds"], "unit": "tet-steps/s", "cores": 1, "kind": "port",
                "sample": f"first {cb['bodies']} bodies of the same {w.name} workload ({cb['tets']} "
                          f"tets), one implicit step from rest each with the plain start "
                          f"({cb['iterations']} CG iterations in all), FP64 ordered-map restatement "
                          f"of fem::simulate, {cb['seconds']:.1f} s on {model} ({ncpu} host cpus, 1 used)",
            }
            result["parity_first_step"] = {
                "bodies": cb["bodies"], "bodies_capped": cb["capped"],
                "rel_displacement_error_max_converged_bodies": cb["err"]["converged"],
                "rel_displacement_error_max_capped_bodies": cb["err"]["capped"],
            }
        else:
            s.set_state(w.coords, np.zeros_like(w.coords))
            s.set_external_forces(w.forces)
            s.simulate(w.dt, True, P)
            xg, _ = s.state()
            log("cpu baseline: one step of the same workload with the FP64 oracle (1 thread)")
            cb = run_cpu_baseline(w, params_kw)
            model, ncpu = cpu_info()
            den = np.linalg.norm(cb["x"] - w.coords, axis=1).max()
            err = np.linalg.norm(xg - cb["x"], axis=1).max() / den if den > 0 else 0.0
            result["cpu_baseline"] = {
                "value": T / cb["seconds"], "unit": "tet-steps/s", "cores": 1, "kind": "port",
                "sample": f"1 implicit step from rest of the same {w.name} workload "
                          f"({cb['iterations']} CG iterations), FP64 ordered-map restatement of "
                          f"fem::simulate, {cb['seconds']:.1f} s on {model} ({ncpu} host cpus, 1 used)",
            }
            result["parity_first_step"] = {"rel_displacement_error": err,
                                           "cg_iterations_cpu": cb["iterations"]}
    s.close()
    del s

    if not args.no_secondary and args.config != "cube64":
        from opentissue_amd import workloads
        w2 = workloads.cube64()
        log(f"secondary: {w2.name}, {w2.num_tets} tets")
        legs2 = tuple(l for l in legs if l in ("learned", "plain", "subspace"))
        s2, _, _, m2 = measure(w2, args, models, -1, legs2)
        sec = {"workload": f"{w2.name}: 1 body, {w2.num_tets} tets, {w2.num_nodes} nodes, dt {w2.dt}, "
                           "CG rel-res 1e-6", "plain": public(m2["plain"]),
               "roofline_plain": roofline(w2, m2["plain"], "cube64"),
               "cg_block": m2["cg_block"], "tile_nodes": m2["tile_nodes"]}
        for leg in legs2:
            if leg != "plain":
                sec[leg] = public(m2[leg])
                sec[f"{leg}_vs_plain"] = m2[leg]["value"] / m2["plain"]["value"]
                sec[f"roofline_{leg}"] = roofline(w2, m2[leg], "cube64")
        result["secondary"] = sec
        s2.close()

    if not args.no_cpu:
        log("cpu baseline: 10x10x40 cantilever, 100 steps in both modes, FP64 oracle (1 thread)")
        result["cpu_baseline_c1"] = cpu_baseline_c1()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
