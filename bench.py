#!/usr/bin/env python
"""Benchmark of the time-stepping hot path on one MI355X.

Workload (default): the 1,038,230-node Velocity-Verlet beam -- 470 x 47 x 47
jittered lattice (seed 1) in a 2.0 x 0.2 x 0.2 m bar, horizon 3.01 dx,
109,555,662 directed bonds, rho 2400, dt 1e-8, a 1e4 N/m^3 body force on the
central strip, supports at both ends.  A step is one launch of the fused
step kernel: break test + bond-force gather over every live bond, force BC,
Velocity-Verlet update.  Inputs are resident in HBM before timing.

Prints one JSON line:
  value          bond-updates/s = live directed bonds x steps / time
  roofline       the step kernel against HBM: algorithmic bytes per launch
                 (4 B per live bond + 276 B per node, the reference's array
                 formats each touched once) / average launch time, measured
                 with HIP events on the library's stream; peak 8 TB/s
  cpu_baseline   the oracle (the reference's Cython kernels restated in C,
                 one core) on a bounded sample of the same workload
  batched        the UQ batch (256 jittered bars of 5k-60k nodes, per-node
                 stiffness fields, Euler dt 1e-5, 500 steps): sample-steps/s
                 of the whole batch in one simulation, its steps captured in
                 one hipGraph, against the same samples run one after another
                 with plain per-step launches (GPU time of back-to-back
                 launches); the batch's launch shape from the autotuner, each
                 candidate timed as a graph run of the whole batch, the
                 sequential runs' from their own fastest shape (timed on 16
                 samples spread over the sizes), with the drop-in's default
                 4 x 128 reported beside it
  composite      the 1M-node composite trilinear Euler-Cromer bar (three
                 bond types, three regimes): step time, bond-updates/s and
                 the HBM fraction with its algorithmic bytes
  adjoint        the stiffness-field adjoint on a 50,820-node jittered bar
                 (105 x 22 x 22, dx = 0.01, the 2560-node bar's loading):
                 2000 Euler steps, checkpoints every 100, loss = squared
                 error of u at 64 sensor nodes against u of a perturbed
                 field (GRF sigma 0.1); forward and backward times and the
                 gradient against central differences (h = 1e-6) on 32
                 random directions, at the configured dt = 1e-5 and at a
                 stable dt = 5e-15

roofline.traffic (HBM bytes per launch) comes from the committed rocprofv3
summary profiles/k_step__<workload>.txt (FETCH_SIZE and WRITE_SIZE passes,
each scaled by the same counter's reading for a plain copy kernel, which
`bench.py --copy-probe` runs); it is null when that summary was made from a
different kernel source or launch shape than the one measured here.
"""
import argparse
import hashlib
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("FP64 bond-updates/s + %HBM peak at 1M nodes; batched vs plain "
          "sample-steps/s")
HBM_PEAK_GBS = 8000.0
REPEATS = 5  # timed regions of K steps; the median is reported
B_NODE = {"euler": 156, "euler_cromer": 252, "velocity_verlet": 276}


def algorithmic_bytes(w, live):
    """Bytes one step must move at least: nlist entries of live bonds (4 B)
    plus the per-node arrays of the integrator, each touched once; n-linear
    models add bond_types + regimes (8 B per live bond)."""
    b = 4 * live + B_NODE[w.integrator] * w.nnodes
    if w.nregimes > 1 or w.nbond_types > 1:
        b += 8 * live
    return b


def kernel_source_hash():
    """sha1 (first 12 hex digits) of the step kernel's source file."""
    src = ROOT / "probabilistic_peridynamics_amd" / "csrc" / "pd_hip.hip"
    return hashlib.sha1(src.read_bytes()).hexdigest()[:12]


def profile_traffic(workload, lpn, tile):
    """HBM bytes per step-kernel launch from profiles/k_step__<workload>.txt
    ("key: value" lines), or None if absent or made from another kernel
    source / launch shape."""
    f = ROOT / "profiles" / f"k_step__{workload}.txt"
    if not f.exists():
        return None
    kv = {}
    for line in f.read_text().splitlines():
        if ":" in line and not line.startswith((" ", "#")):
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    try:
        if (kv.get("kernel_source_sha1") != kernel_source_hash()
                or int(kv.get("lanes_per_node", -1)) != lpn
                or int(kv.get("tile_nodes", -1)) != tile):
            return None
        return float(kv["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def copy_probe(mib):
    """A plain copy kernel of known size (read mib MiB, write mib MiB), five
    launches, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8,
                      "copy_probe_bytes_written": n * 8, "launches": 5}),
          flush=True)
    return 0


def cpu_multicore(w, budget_s):
    """The oracle's per-node (gather) formulation of the same step -- the
    slot-layout bond force of cl/peridynamics.cl restated with the Cython
    stretch, then the integrator's update -- on all host cores given to this
    process (at most 16), whole steps until budget_s has passed."""
    import oracle
    L = oracle.lib()
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    L.oracle_set_threads(cores)
    n, k = w.nnodes, w.max_neighbours
    nlist = np.array(w.nlist, dtype=np.int32)
    u, ud, udd = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    force, body = np.zeros((n, 3)), np.zeros((n, 3))
    d, f = w.magnitudes(10000)
    live = int(w.n_neigh.sum())
    D, I = oracle._d, oracle._i
    upd = (L.oracle_update_velocity_verlet
           if w.integrator == "velocity_verlet"
           else L.oracle_update_euler_cromer)
    steps = 0
    t0 = time.perf_counter()
    while True:
        L.oracle_slot_bond_force(
            n, k, D(u), D(w.coords), I(nlist), D(w.volume),
            float(w.bond_stiffness), float(w.critical_stretch), 0, None,
            None, None, None, None, None, None, 1, I(w.force_bc_types),
            D(w.force_bc_values), f[steps], D(force), D(body), 1)
        upd(n, D(force), D(u), D(ud), D(udd), I(w.bc_types),
            D(w.bc_values), D(w.densities), d[steps], w.damping, w.dt)
        steps += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    elapsed = time.perf_counter() - t0
    L.oracle_set_threads(1)
    return {
        "value": live * steps / elapsed, "unit": "bond-updates/s",
        "cores": cores, "kind": "port",
        "sample": (f"{steps} full steps of the {n}-node workload "
                   f"({elapsed:.1f} s): the per-node slot-layout bond force "
                   "and the "
                   f"{w.integrator} update of cl/*.cl restated in C "
                   "(gcc -O3, OpenMP)"),
    }


def uq_bench(nsamples, steps, tune_steps):
    """Batched vs sequential sample-steps/s on the UQ batch."""
    from probabilistic_peridynamics_amd import autotune, batch, configs
    t0 = time.perf_counter()
    ws = configs.uq_samples(nsamples)
    nodes = sum(w.nnodes for w in ws)
    log(f"[bench] uq: {len(ws)} samples, {nodes} nodes, built in "
        f"{time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    b = batch.SampleBatch(ws)
    log(f"[bench] uq batch set up in {time.perf_counter() - t0:.1f}s")
    # the whole batch, each candidate as a graph run (the timed form)
    t0 = time.perf_counter()
    best, table = autotune.autotune_batch(b, steps=tune_steps, graph=True)
    log(f"[bench] uq autotune ({time.perf_counter() - t0:.1f}s): best {best} "
        + " ".join(f"{c}:{ms:.2f}ms" for c, ms in table))
    log(f"[bench] uq: {b.integrator.tile_stats()['ntiles']} tiles")
    ms_batched = b.run(steps, graph=True)
    live_end = b.integrator.live_bonds()
    del b

    def sequential(shape, subset, nsteps):
        """GPU ms of the samples run one after another, each in its own
        simulation with back-to-back plain launches (pd_run)."""
        tot = 0.0
        for w in subset:
            integ = batch.solo_integrator(w, lanes_per_node=shape[0],
                                          tile_nodes=shape[1])
            d, f = w.magnitudes(nsteps)
            tot += integ.timed_run(nsteps, d, f)
            del integ
        return tot

    # the sequential runs get their own launch shape: small meshes fill the
    # GPU better with more, smaller workgroups than the batch does
    default = (4, 128)
    seq_cands = [default, (8, 64), (4, 64), (2, 128), (2, 256)]
    spread = ws[::max(1, len(ws) // 16)]
    seq_table = [(c, sequential(c, spread, 20)) for c in seq_cands]
    seq_best = min(seq_table, key=lambda e: e[1])[0]
    log("[bench] uq sequential shapes (16 samples, 20 steps): "
        + " ".join(f"{c}:{ms:.1f}ms" for c, ms in seq_table))
    ms_seq = sequential(seq_best, ws, steps)
    ms_seq_default = (ms_seq if seq_best == default
                      else sequential(default, ws, steps))
    ss_b = len(ws) * steps / (ms_batched / 1e3)
    ss_s = len(ws) * steps / (ms_seq / 1e3)
    ss_d = len(ws) * steps / (ms_seq_default / 1e3)
    log(f"[bench] uq: batched {ms_batched:.1f} ms, sequential {ms_seq:.1f} "
        f"ms at {seq_best} ({ms_seq_default:.1f} ms at {default}) for "
        f"{steps} steps")
    return {
        "workload": f"uq_{len(ws)}_samples_{nodes}_nodes",
        "samples": len(ws), "nodes": nodes, "steps": steps,
        "integrator": "euler", "dt": ws[0].dt,
        "batched_sample_steps_per_s": ss_b,
        "sequential_sample_steps_per_s": ss_s,
        "speedup": ss_b / ss_s,
        "batched_ms": ms_batched, "sequential_ms": ms_seq,
        "sequential_launch_shape": {"lanes_per_node": seq_best[0],
                                    "tile_nodes": seq_best[1]},
        "sequential_shapes": [[list(c), ms] for c, ms in seq_table],
        "sequential_default_shape_sample_steps_per_s": ss_d,
        "speedup_vs_default_shape": ss_b / ss_d,
        "live_bonds_start": int(sum(w.n_neigh.sum() for w in ws)),
        "live_bonds_end": live_end,
        "launch_shape": {"lanes_per_node": best[0], "tile_nodes": best[1]},
        "autotune": [[list(c), ms] for c, ms in table],
        "graph": True,
    }


def composite_bench(steps, warmup=3):
    """Step time of the composite n-linear Euler-Cromer workload."""
    from probabilistic_peridynamics_amd import configs
    w = configs.composite_ec()
    integ = configs.make_integrator(w)
    dm, fm = w.magnitudes(warmup + steps)
    integ.run(warmup, dm[:warmup], fm[:warmup])
    integ.synchronize()
    live0 = integ.live_bonds()
    ms = integ.timed_run(steps, dm[warmup:], fm[warmup:])
    live1 = integ.live_bonds()
    per = ms / 1e3 / steps
    live = 0.5 * (live0 + live1)
    B = algorithmic_bytes(w, live)
    log(f"[bench] composite: {per * 1e6:.1f} us/step, live {live0}->{live1}")
    return {
        "workload": w.name, "nodes": w.nnodes, "live_bonds": live,
        "integrator": w.integrator, "steps": steps,
        "ms_per_step": per * 1e3, "bond_updates_per_s": live / per,
        "algorithmic_bytes_per_launch": B,
        "hbm_frac": B / per / 1e9 / HBM_PEAK_GBS,
    }


def exact_bench(steps, warmup=5):
    """The Cython path's Euler on the 1,038,230-node beam (compact layout):
    ms per step of the exact-order step (each node's bonds summed in
    peridynamics.pyx's order, the default of integrators.Euler) and of the
    fast-order step."""
    from probabilistic_peridynamics_amd import configs
    w = configs.beam_euler()
    integ = configs.make_integrator(w)
    dm, fm = w.magnitudes(warmup + 2 * (2 + REPEATS * steps))
    out = {"workload": w.name, "nodes": w.nnodes,
           "live_bonds": w.live_bonds(), "integrator": "euler (compact)"}
    at = warmup
    integ.run(warmup, dm[:at], fm[:at])
    for mode in (True, False):
        integ.set_exact_order(mode)
        integ.run(2, dm[at:at + 2], fm[at:at + 2])
        at += 2
        reps = []
        for r in range(REPEATS):
            reps.append(integ.timed_run(steps, dm[at:at + steps],
                                        fm[at:at + steps]) / steps)
            at += steps
        out["exact_ms_per_step" if mode else "fast_ms_per_step"] = float(
            np.median(reps))
    out["exact_over_fast"] = out["exact_ms_per_step"] / out["fast_ms_per_step"]
    log(f"[bench] exact order: {out['exact_ms_per_step'] * 1e3:.1f} us/step, "
        f"fast order {out['fast_ms_per_step'] * 1e3:.1f} us/step")
    return out


def adjoint_bench(steps, every, ndirs, dt):
    """Forward/backward timing and FD agreement of the adjoint."""
    from probabilistic_peridynamics_amd import adjoint, configs
    w = configs.uq_sample((105, 22, 22), 4)
    w.dt = dt
    rng = np.random.default_rng(4)
    g_true = configs.gaussian_random_field((105, 22, 22), 0.01, 0.05, 0.1,
                                           rng).reshape(-1)
    sensors = rng.choice(w.nnodes, 64, replace=False)
    solver = adjoint.StiffnessAdjoint(w, steps, checkpoint_every=every)
    target = solver.forward(g_true)[sensors]
    g = np.zeros(w.nnodes)

    def loss(gg):
        u = solver.forward(gg)
        return float(((u[sensors] - target) ** 2).sum()), u

    t0 = time.perf_counter()
    L, u = loss(g)
    t_fwd = time.perf_counter() - t0
    dL = np.zeros_like(u)
    dL[sensors] = 2.0 * (u[sensors] - target)
    t0 = time.perf_counter()
    grad = solver.backward(dL)
    t_bwd = time.perf_counter() - t0
    _, dmg, _ = solver.state()
    vs = [rng.standard_normal(w.nnodes) for _ in range(ndirs)]
    fd_err = {}
    for h in (1e-6, 1e-4):
        errs = []
        for v in vs:
            fd = (loss(g + h * v)[0] - loss(g - h * v)[0]) / (2 * h)
            an = float(grad @ v)
            errs.append(abs(an - fd) / max(abs(fd), 1e-300))
        fd_err[h] = np.array(errs)
    errs = fd_err[1e-6]
    finite = bool(np.all(np.isfinite(grad)))
    log(f"[bench] adjoint dt={dt}: L={L:.4e} fwd {t_fwd:.2f}s bwd "
        f"{t_bwd:.2f}s fd(h=1e-6) rel err max {errs.max():.3e} median "
        f"{np.median(errs):.3e}; fd(h=1e-4) max {fd_err[1e-4].max():.3e} "
        f"finite {finite}")
    return {
        "workload": f"adjoint_bar_{w.nnodes}", "nodes": w.nnodes,
        "steps": steps, "checkpoint_every": every, "dt": dt,
        "sensors": 64, "loss": L, "forward_s": t_fwd, "backward_s": t_bwd,
        "gradient_finite": finite,
        "fd_directions": ndirs, "fd_rel_err_max": float(errs.max()),
        "fd_rel_err_median": float(np.median(errs)),
        "fd_h1e-4_rel_err_max": float(fd_err[1e-4].max()),
        "fd_h1e-4_rel_err_median": float(np.median(fd_err[1e-4])),
        "broken_fraction": float(np.mean(dmg)),
    }


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(w, budget_s):
    """Time the oracle on this workload: break_bonds + bond_force
    (peridynamics.pyx order, one thread) and the integrator's update, whole
    steps until budget_s has passed (at least one)."""
    import oracle
    oracle.lib().oracle_set_threads(1)
    n = w.nnodes
    coords = w.coords
    nlist = np.array(w.nlist, dtype=np.int32)
    n_neigh = np.array(w.n_neigh, dtype=np.int32)
    u = np.zeros((n, 3))
    ud = np.zeros((n, 3))
    udd = np.zeros((n, 3))
    L = oracle.lib()
    d, f = w.magnitudes(10000)
    kind = w.integrator
    steps = 0
    live = int(n_neigh.sum())
    t0 = time.perf_counter()
    while True:
        oracle.break_bonds(coords + u, coords, nlist, n_neigh,
                           w.critical_stretch)
        force = oracle.bond_force(coords + u, coords, nlist, n_neigh,
                                  w.volume, w.bond_stiffness,
                                  w.force_bc_values, w.force_bc_types,
                                  f[steps])
        if kind == "euler":
            oracle.update_displacement(u, w.bc_values, w.bc_types, force,
                                       d[steps], w.dt)
        else:
            fn = (L.oracle_update_velocity_verlet
                  if kind == "velocity_verlet"
                  else L.oracle_update_euler_cromer)
            fn(n, oracle._d(force), oracle._d(u), oracle._d(ud),
               oracle._d(udd), oracle._i(w.bc_types),
               oracle._d(w.bc_values), oracle._d(w.densities), d[steps],
               w.damping, w.dt)
        steps += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    elapsed = time.perf_counter() - t0
    return {
        "value": live * steps / elapsed,
        "unit": "bond-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{steps} full steps of the {n}-node workload "
                   f"({elapsed:.1f} s): peridynamics.pyx break_bonds + "
                   "bond_force restated in C (gcc -O3), plus the "
                   f"{kind} update"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="beam_vv",
                    choices=["beam_vv", "composite_ec", "bar_euler"])
    ap.add_argument("--lanes-per-node", type=int, default=4)
    ap.add_argument("--tile-nodes", type=int, default=128)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU baseline budget (0 to skip)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the counter-calibration copy kernel")
    ap.add_argument("--uq-samples", type=int, default=256,
                    help="samples of the batched UQ comparison (0 to skip)")
    ap.add_argument("--uq-steps", type=int, default=500)
    ap.add_argument("--uq-tune-steps", type=int, default=10,
                    help="steps per autotuner candidate (graph runs of the "
                    "whole batch)")
    ap.add_argument("--composite-steps", type=int, default=50,
                    help="steps of the composite workload (0 to skip)")
    ap.add_argument("--exact-steps", type=int, default=20,
                    help="steps per timed region of the exact-order "
                    "measurement (0 to skip)")
    ap.add_argument("--adjoint-steps", type=int, default=2000,
                    help="steps of the adjoint check (0 to skip)")
    ap.add_argument("--adjoint-dirs", type=int, default=32)
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU "
                          "sharding is a design (see the design notes), "
                          "not an implemented path", "n_gpus": args.gpus}))
        return 2

    from probabilistic_peridynamics_amd import _lib, configs

    if _lib.device_count() < 1:
        raise SystemExit("bench.py needs a HIP GPU")
    t_setup = time.perf_counter()
    w = getattr(configs, args.workload)()
    integ = configs.make_integrator(w, lanes_per_node=args.lanes_per_node,
                                    tile_nodes=args.tile_nodes)
    tiles = integ.tile_stats()
    setup_s = time.perf_counter() - t_setup
    print(f"[bench] {w.name}: N={w.nnodes} L={w.live_bonds()} "
          f"K={w.max_neighbours} setup {setup_s:.1f}s tiles {tiles}",
          file=sys.stderr, flush=True)

    # W untimed steps, then REPEATS timed regions of exactly K steps each
    # (HIP events on the library's stream, synchronized on both sides); the
    # median region is reported and all of them are listed
    dm, fm = w.magnitudes(args.warmup + REPEATS * args.steps)
    integ.run(args.warmup, dm[:args.warmup], fm[:args.warmup])
    integ.synchronize()
    live0 = integ.live_bonds()
    reps = []
    t0 = time.perf_counter()
    for r in range(REPEATS):
        lo = args.warmup + r * args.steps
        integ.synchronize()
        reps.append(integ.timed_run(args.steps, dm[lo:lo + args.steps],
                                    fm[lo:lo + args.steps]))
        integ.synchronize()
    wall = time.perf_counter() - t0
    live1 = integ.live_bonds()
    live = 0.5 * (live0 + live1)
    ms = float(np.median(reps))
    per_step_s = ms / 1e3 / args.steps
    value = live / per_step_s
    B = algorithmic_bytes(w, live)
    achieved = B / per_step_s / 1e9
    print(f"[bench] {REPEATS} x {args.steps} steps: "
          + " ".join(f"{x:.3f}" for x in reps)
          + f" ms (events, median {ms:.3f}), {wall*1e3:.1f} ms wall, live "
          f"{live0}->{live1}", file=sys.stderr, flush=True)

    del integ
    uq = None
    if args.uq_samples > 0:
        uq = uq_bench(args.uq_samples, args.uq_steps, args.uq_tune_steps)

    comp = None
    if args.composite_steps > 0:
        comp = composite_bench(args.composite_steps)

    exact = None
    if args.exact_steps > 0:
        exact = exact_bench(args.exact_steps)

    adj = None
    if args.adjoint_steps > 0:
        adj = [adjoint_bench(args.adjoint_steps, 100, args.adjoint_dirs, dt)
               for dt in (1e-5, 5e-15)]

    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(w, args.cpu_seconds)
        cpu["threads_available"] = os.cpu_count()
        if w.integrator != "euler":
            cpu["multicore"] = cpu_multicore(w, args.cpu_seconds)

    name, arch = _lib.device_name(0)
    traffic = profile_traffic(w.name, args.lanes_per_node, args.tile_nodes)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "bond-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": per_step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": w.name,
            "nodes": w.nnodes,
            "live_bonds": live,
            "max_neighbours": w.max_neighbours,
            "integrator": w.integrator,
            "lanes_per_node": args.lanes_per_node,
            "tile_nodes": args.tile_nodes,
            "tiles": tiles["ntiles"],
            "halo_per_node": tiles["halo_total"] / w.nnodes,
            "device": f"{name} ({arch})",
            "kernel_source_sha1": kernel_source_hash(),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_frac": (None if traffic is None else
                             traffic / per_step_s / 1e9 / HBM_PEAK_GBS),
            "algorithmic_bytes_per_launch": B,
            "launch_us": per_step_s * 1e6,
            "repeats_ms_per_step": [x / args.steps for x in reps],
        },
        "cpu_baseline": cpu,
        "batched": uq,
        "composite": comp,
        "exact_order": exact,
        "adjoint": adj,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
