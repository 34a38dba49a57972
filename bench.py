#!/usr/bin/env python3
"""Throughput of the corotated tetrahedral FEM step on one MI355X.

Default workload: the 64^3-cube lattice (1,310,720 tets, 274,625 nodes),
interior nodes jittered +/-0.2 h, E log-uniform per tet in [1e5, 1e6],
nu 0.3, face x = 0 fixed, 1e3 Pa traction on the opposite face, dt 0.005,
Rayleigh alpha 0.1 / beta 0.001, CG run until |b - A v| <= 1e-6 |b| (a
safety cap of 5,000 iterations; `--cap 200` gives the reference case's cap,
under which most steps of this mesh stop short of 1e-6).
A step is one full implicit step (rotations + elastic forces + right-hand
side + CG + position update) over the whole mesh, inputs resident in HBM.
Setup (not timed): the bodies are built on the device, the layout autotuner
picks the tile target and CG block size on throw-away solvers
(`--tile-nodes 0` keeps the default tiling and tunes the block size only),
then the solver is created.

Output: one JSON line.  `value` is tet-steps/s over exactly --steps timed
steps from rest with the plain CG start (the reference's v^n); `learned`
repeats the timed steps with the learned CG start (network trained first, on
2,000 plain-solver steps of the 10x10x40 cantilever under seeded random
tractions; training is not timed) and gives the learned-vs-plain A/B on
held-out cantilever episodes with CG run to rel-res 1e-6 uncapped.
`roofline` prices the dominant kernels (one k_cg_iter launch per CG
iteration, plus k_finish) by their algorithmic bytes (100 B/tet + 195 B/node
per CG iteration, FP64 v and r, plus 36 B/node for the position update)
against the MI355X HBM peak; `cpu_baseline` times the FP64 CPU restatement of the reference step
(one thread) on a bounded sample of the same workload.

    python bench.py [--steps 200] [--warmup 5] [--config cube64|cantilever|bodies|combined]
    python bench.py --copy-probe     # plain device copy of known size, for
                                     # scaling rocprofv3 FETCH/WRITE counters
"""
import argparse
import json
import os
import platform
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


def run_cpu_baseline(w, params_kw, seconds_budget):
    """One implicit step of the same workload from rest with the FP64 oracle."""
    import oracle
    t_init0 = time.perf_counter()
    o = oracle.Oracle(w.coords, w.tets, w.fixed)
    o.init_materials(w.young, w.poisson, w.density)
    o.set_external_forces(w.forces)
    t_init = time.perf_counter() - t_init0
    P = oracle.config_params(**params_kw)
    t0 = time.perf_counter()
    st = o.simulate(w.dt, True, P)
    t = time.perf_counter() - t0
    x, v = o.state()
    return dict(seconds=t, init_seconds=t_init, iterations=st.iterations, x=x, v=v)


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


def train_warmstart_model():
    """Network for the learned CG start: 2,000 plain-solver steps (20 episodes
    of 100) of the 10x10x40 cantilever under seeded random tractions, then
    Adam on the residual loss.  Returns (model, solver, workload, info)."""
    from opentissue_amd import fem, warmstart, workloads
    w = workloads.cantilever()
    s = warmstart.make_solver(w)
    P = fem.config_params()
    model = fem.Model(seed=0)
    t0 = time.perf_counter()
    loads, _ = warmstart.random_loads(w, 20, seed=1)
    n = warmstart.record(s, model, w, loads, 100, P)
    t1 = time.perf_counter()
    loss = model.train_residual(s, w.dt, P, steps=250, lr=1e-4, seed=0)
    t2 = time.perf_counter()
    info = dict(samples=int(n), record_seconds=t1 - t0, train_seconds=t2 - t1, adam_steps=250,
                lr=1e-4, nodes_per_batch=int(s.n_free * max(1, round(65536 / s.n_free))),
                final_loss=float(np.mean(loss[-50:])))
    return model, s, w, info


def learned_ab(model, s, w):
    """Held-out cantilever episodes (seed 1001, 3 x 100 steps from rest): plain
    vs learned start, CG to rel-res 1e-6 with cap 3000 (every step converges)."""
    from opentissue_amd import fem, warmstart
    U = fem.config_params(cap=3000)
    loads, _ = warmstart.random_loads(w, 3, seed=1001)
    plain = warmstart.run_episodes(s, w, loads, 100, U, learned=False)
    learned = warmstart.run_episodes(s, w, loads, 100, U, learned=True, model=model)
    return {
        "workload": f"{w.name} ({w.num_tets} tets), 3 held-out episodes x 100 steps, "
                    "random traction 0.5e3..2e3 Pa, CG rel-res 1e-6 cap 3000",
        "plain": {"tet_steps_per_s": plain["tet_steps_per_s"],
                  "cg_iterations_per_step": plain["mean_iterations"]},
        "learned": {"tet_steps_per_s": learned["tet_steps_per_s"],
                    "cg_iterations_per_step": learned["mean_iterations"]},
        "speedup": learned["tet_steps_per_s"] / plain["tet_steps_per_s"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="cube64")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="time budget of the multi-body CPU sample")
    ap.add_argument("--tile-nodes", type=int, default=-1,
                    help="owned nodes per tile (0: the default tiling; -1: the layout autotuner "
                         "picks the tile target and the CG block size)")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm start legs")
    ap.add_argument("--cap", type=int, default=5000, help="CG iteration cap (safety net)")
    ap.add_argument("--cg-block", type=int, default=0, choices=(0, 256, 512, 1024),
                    help="threads per tile in the CG kernels (0: autotune)")
    ap.add_argument("--deterministic", action="store_true",
                    help="per-node sums in slot order (same bits on every run)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU "
              "sharding of this path is a design note, not a measured configuration.",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return

    from opentissue_amd import fem

    t_setup = time.perf_counter()
    w = build_workload(args.config)
    t_built = time.perf_counter()
    log(f"workload {w.name}: {w.num_bodies} bodies, {w.num_tets} tets, {w.num_nodes} nodes "
        f"(built in {t_built - t_setup:.2f} s)")
    params_kw = dict(alpha=0.1, beta=0.001, tol=1e-6, cap=args.cap)
    P = fem.config_params(**params_kw)
    tile_nodes, tuned_block = args.tile_nodes, 0
    if tile_nodes < 0:
        tile_nodes, tuned_block, _ = fem.autotune_layout(
            w.coords, w.tets, w.fixed, w.young, w.poisson, w.density, w.forces, w.dt, P, steps=2,
            body_node_offsets=w.body_node_offsets, body_tet_offsets=w.body_tet_offsets)
        log(f"layout autotuner: target_tile_nodes {tile_nodes} ({'default tiling' if tile_nodes == 0 else 'owned nodes per tile'}), "
            f"CG block {tuned_block} ({time.perf_counter() - t_built:.1f} s)")
    t_create = time.perf_counter()
    s = fem.Solver(w.coords, w.tets, w.fixed, body_node_offsets=w.body_node_offsets,
                   body_tet_offsets=w.body_tet_offsets, target_tile_nodes=tile_nodes)
    log(f"otfem_create (tile layout) {time.perf_counter() - t_create:.2f} s")
    s.init_materials(w.young, w.poisson, w.density)
    s.set_external_forces(w.forces)
    if args.deterministic:
        s.set_deterministic(True)
    layout = s.layout_info()
    if args.cg_block or tuned_block:
        best_block = args.cg_block or tuned_block
        s.set_kernel_config(cg_block=best_block)
        log(f"CG block size: {best_block} threads per tile ({layout['tiles']} tiles)")
    else:
        best_block = s.autotune(w.dt, True, P, steps=2)
        log(f"autotuned CG block size: {best_block} threads per tile ({layout['tiles']} tiles)")

    # warmup, then back to rest so the timed steps are steps 1..K from rest
    for _ in range(args.warmup):
        s.step_async(w.dt, True, P)
    s.synchronize()
    zeros = np.zeros_like(w.coords)
    s.set_state(w.coords, zeros)
    s.iteration_totals(reset=True)
    s.unconverged_totals(reset=True)
    s.set_timing(True)
    s.synchronize()

    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.step_async(w.dt, True, P)
    st = s.synchronize()
    t1 = time.perf_counter()
    unconverged = int(s.unconverged_totals().sum())
    elem_ms, cg_ms, nsteps = s.timing()
    iters = s.iteration_totals()
    wall = t1 - t0
    T, N = w.num_tets, w.num_nodes
    value = T * args.steps / wall
    total_iters = int(iters.sum())
    # algorithmic bytes of all CG launches in the timed region
    if w.num_bodies == 1:
        cg_bytes = total_iters * (BYTES_PER_TET_ITER * T + BYTES_PER_NODE_ITER * N)
    else:
        bt = np.diff(w.body_tet_offsets)
        bn = np.diff(w.body_node_offsets)
        cg_bytes = float(np.sum(iters * (BYTES_PER_TET_ITER * bt + BYTES_PER_NODE_ITER * bn)))
    cg_bytes += args.steps * BYTES_PER_NODE_POSUPDATE * N
    achieved = cg_bytes / (cg_ms / 1e3) / 1e9 if cg_ms > 0 else 0.0
    log(f"timed {args.steps} steps in {wall:.3f} s; element {elem_ms:.2f} ms, CG {cg_ms:.2f} ms; "
        f"CG iterations {total_iters} ({total_iters / max(1, args.steps) / w.num_bodies:.1f} per step per body)")

    # figures of the committed rocprofv3 summaries of this workload, if any
    traffic = profile_avg_us = None
    tpath = os.path.join(ROOT, "profiles", f"k_cg_iter_{args.config}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    kpath = os.path.join(ROOT, "profiles", f"k_cg_iter_{args.config}_kernel_trace.json")
    if os.path.exists(kpath):
        with open(kpath) as f:
            profile_avg_us = json.load(f).get("avg_us_per_working_launch")
    per_launch = {}
    if w.num_bodies == 1 and total_iters > 0:
        # one k_cg_iter launch is one CG iteration over the whole mesh
        per_launch = {
            "algorithmic_bytes_per_launch": BYTES_PER_TET_ITER * T + BYTES_PER_NODE_ITER * N,
            "avg_launch_us": 1e3 * cg_ms / total_iters,
            "avg_launch_note": "HIP events over the timed CG work divided by the CG iterations: one "
                               "k_cg_iter launch plus its share of k_finish (position update), the "
                               "early-exit launches of the last chunk and launch gaps",
        }

    result = {
        "metric": "tet-steps/s at CG rel-res 1e-6, learned vs plain; CG HBM GB/s % of peak",
        "value": value,
        "unit": "tet-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {
            "workload": f"{w.name}: {w.num_bodies} body/bodies, {T} tets, {N} nodes, "
                        f"dt {w.dt}, CG rel-res 1e-6 (cap {args.cap}), alpha 0.1 beta 0.001, "
                        "warm start v^i",
            "tets": T, "nodes": N, "bodies": w.num_bodies,
            "tiles": layout["tiles"], "slots": layout["slots"], "cg_block": best_block,
            "tile_nodes": tile_nodes, "layout_autotuned": args.tile_nodes < 0,
            "deterministic": bool(args.deterministic),
        },
        "cg_iterations_per_step": total_iters / args.steps / w.num_bodies,
        "cg_body_steps_unconverged": unconverged,
        "cg_body_steps": args.steps * w.num_bodies,
        "kernel_ms_per_step": {"element": elem_ms / max(1, nsteps), "cg": cg_ms / max(1, nsteps)},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_note": "HBM bytes per working k_cg_iter launch (rocprofv3 FETCH_SIZE / "
                            "WRITE_SIZE scaled by a copy of known size)" if traffic else None,
            "kernel": "k_cg_iter (one launch per CG iteration) + k_finish",
            "algorithmic_bytes_per_step": cg_bytes / args.steps,
            "cg_ms_per_step": cg_ms / max(1, nsteps),
            **per_launch,
            "profile_avg_launch_us": profile_avg_us,
        },
        "cpu_baseline": None,
    }

    # A caller that keeps the node arrays in host memory (the fem.hpp drop-in)
    # uploads x, v and f_ext and reads x, v back every step.  The same first
    # steps from rest, device-resident and through host buffers; never `value`.
    k2 = min(args.steps, 20)
    s.set_timing(False)
    s.set_state(w.coords, zeros)
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(k2):
        s.step_async(w.dt, True, P)
    s.synchronize()
    dev_s = time.perf_counter() - t0
    xh, vh = w.coords.copy(), zeros.copy()
    t0 = time.perf_counter()
    for _ in range(k2):
        s.set_state(xh, vh)
        s.set_external_forces(w.forces)
        s.simulate(w.dt, True, P)
        xh, vh = s.state()
    host_s = time.perf_counter() - t0
    result["host_buffers"] = {
        "value": T * k2 / host_s,
        "unit": "tet-steps/s",
        "device_resident_same_steps": T * k2 / dev_s,
        "steps": k2,
        "note": "steps 1..%d from rest with x, v, f_ext copied from pageable host arrays and x, v "
                "copied back every step (%.1f MB up, %.1f MB down), against the same steps "
                "device-resident" % (k2, 9 * 8 * N / 1e6, 6 * 8 * N / 1e6),
    }
    log(f"host-buffer caller: {T * k2 / host_s:.4g} tet-steps/s vs {T * k2 / dev_s:.4g} "
        f"device-resident over the first {k2} steps")

    if not args.no_learned:
        from opentissue_amd import fem as F
        log("learned warm start: recording 2,000 cantilever steps and training (not timed)")
        model, cs, cw, info = train_warmstart_model()
        ab = learned_ab(model, cs, cw)
        cs.close()
        # the timed steps again, CG started from the network
        PL = F.Params.from_buffer_copy(P)
        PL.warm_start = F.WARM_LEARNED
        s.set_warmstart_model(model)
        s.set_state(w.coords, zeros)
        s.iteration_totals(reset=True)
        s.unconverged_totals(reset=True)
        s.set_timing(True)
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            s.step_async(w.dt, True, PL)
        stl = s.synchronize()
        wl = time.perf_counter() - t0
        itl = s.iteration_totals()
        unl = int(s.unconverged_totals().sum())
        el_ms, cgl_ms, ws_ms, nl = s.timing4()
        # network roofline: nominal 2 * 34,688 MAC per node per forward pass
        net_flops = 2.0 * 34688 * N
        net_ms = ws_ms / max(1, nl)
        net_tflops = net_flops / (net_ms / 1e3) / 1e12 if net_ms > 0 else 0.0
        s.set_warmstart_model(None)
        vl = T * args.steps / wl
        log(f"learned start: {vl:.4g} tet-steps/s ({vl / value:.3f}x plain), "
            f"{itl.sum() / args.steps / w.num_bodies:.1f} CG iterations per step; "
            f"A/B on held-out cantilever episodes (uncapped): {ab['speedup']:.3f}x")
        result["learned"] = {
            "value": vl,
            "unit": "tet-steps/s",
            "vs_plain": vl / value,
            "cg_iterations_per_step": float(itl.sum()) / args.steps / w.num_bodies,
            "cg_body_steps_unconverged": unl,
            "kernel_ms_per_step": {"network": net_ms, "element": el_ms / max(1, nl),
                                   "cg": cgl_ms / max(1, nl)},
            "network_roofline": {"bound": "mfma", "kernel": "k_ws_predict",
                                 "achieved": net_tflops, "peak": 2500.0, "unit": "TFLOP/s",
                                 "frac": net_tflops / 2500.0,
                                 "algorithmic_flops_per_launch": net_flops,
                                 "avg_launch_ms": net_ms,
                                 "note": "nominal BF16 flops; the kernel issues 2.2x as many "
                                         "(inputs and activations as BF16 hi + lo pairs, "
                                         "3-output layer padded to 32 rows)"},
            "flags": int(stl.flags),
            "training": info,
            "ab_uncapped": ab,
        }

    if not args.no_cpu and w.num_bodies > 1:
        # bounded sample: bodies of the same workload, one step from rest each
        s.set_state(w.coords, zeros)
        s.simulate(w.dt, True, P)
        xg, _ = s.state()
        log(f"cpu baseline: bodies of the same workload, one step each, FP64 oracle (1 thread), "
            f"{args.cpu_seconds:.0f} s budget")
        cb = run_cpu_baseline_bodies(w, params_kw, args.cpu_seconds, xg)
        model, ncpu = cpu_info()
        result["cpu_baseline"] = {
            "value": cb["tets"] / cb["seconds"],
            "unit": "tet-steps/s",
            "cores": 1,
            "kind": "port",
            "sample": f"first {cb['bodies']} bodies of the same {w.name} workload ({cb['tets']} tets), "
                      f"one implicit step from rest each ({cb['iterations']} CG iterations in all), "
                      f"FP64 ordered-map restatement of fem::simulate, {cb['seconds']:.1f} s on {model} "
                      f"({ncpu} host cpus, 1 used)",
        }
        result["parity_first_step"] = {
            "bodies": cb["bodies"],
            "bodies_capped": cb["capped"],
            "rel_displacement_error_max_converged_bodies": cb["err"]["converged"],
            "rel_displacement_error_max_capped_bodies": cb["err"]["capped"],
        }
    elif not args.no_cpu:
        # bounded sample: one step from rest of the same workload; the GPU's
        # first step from the same state is the parity check at full size
        s2 = fem.Solver(w.coords, w.tets, w.fixed, body_node_offsets=w.body_node_offsets,
                        body_tet_offsets=w.body_tet_offsets, target_tile_nodes=tile_nodes)
        s2.init_materials(w.young, w.poisson, w.density)
        s2.set_external_forces(w.forces)
        s2.simulate(w.dt, True, P)
        xg, _ = s2.state()
        s2.close()
        log("cpu baseline: one step of the same workload with the FP64 oracle (1 thread)")
        cb = run_cpu_baseline(w, params_kw, args.cpu_seconds)
        model, ncpu = cpu_info()
        den = np.linalg.norm(cb["x"] - w.coords, axis=1).max()
        err = np.linalg.norm(xg - cb["x"], axis=1).max() / den if den > 0 else 0.0
        result["cpu_baseline"] = {
            "value": T / cb["seconds"],
            "unit": "tet-steps/s",
            "cores": 1,
            "kind": "port",
            "sample": f"1 implicit step from rest of the same {w.name} workload "
                      f"({cb['iterations']} CG iterations), FP64 ordered-map restatement of "
                      f"fem::simulate, {cb['seconds']:.1f} s on {model} ({ncpu} host cpus, 1 used)",
        }
        result["parity_first_step"] = {"rel_displacement_error": err,
                                       "cg_iterations_cpu": cb["iterations"]}
    print(json.dumps(result))


if __name__ == "__main__":
    main()
