#!/usr/bin/env python3
"""Plain-step throughput of the DEM engine on one MI355X.

Workload: silo discharge, 1,048,576 polydisperse spheres (r ~ U[2, 4] mm,
rho 2600), Hertz-Mindlin (E 1e7 Pa, nu 0.3, e 0.5, mu 0.5) with tangential
spring history, cylindrical silo R 25 cm over a 30-degree conical hopper with
a 6 cm outlet, cell 2*rmax, dt 5e-6 s.  Setup (untimed): the seeded random
non-overlapping fill, settled with the outlet closed at a larger time step
until the fastest particle moves below --quiet-vmax, then the outlet opens
(Device_OpenHatch) and the time step returns to 5e-6.  W warm-up steps, then
exactly K timed steps of the full step (sort, gather, cell table, fused
contact/force/wall/integrate kernel), bracketed by stream synchronisation.

Metric: particle-steps/s = live particles x K / elapsed.  One JSON line on
stdout; progress on stderr.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "softx-d-15-00085_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_device(bdem, deck):
    dev = bdem.Device(deck["wobj"], deck["pobj"], deck["sim"])
    dev.set_ext(deck["ext"])
    dev.set_radii(deck["radius"], deck["mass"])
    dev.set_state(deck["pos"], deck["ptype"], deck["pid"], vel=deck["vel"], avel=deck["avel"])
    return dev


def step_kernel_bytes(n, k, ncells):
    """Algorithmic bytes of one fused-step launch (history law), per the
    stage I/O: reads state 48 + sorted key 4 + packed spring counts 4 +
    springs 16k; writes state 48 + key 4 + changed flag 1 + spring counts 4 +
    springs 16k; plus the cell table 8C once."""
    per = 48 + 4 + 4 + 16 * k + 48 + 4 + 1 + 4 + 16 * k
    return n * per + 8 * ncells


def step_total_bytes(n, k, ncells):
    """Whole-step algorithmic bytes: 320 + 8k + 8C/N + 64k per particle
    (history law), the per-particle-step figure of the measurement plan."""
    return n * (320 + 8 * k + 64 * k) + 8 * ncells


def cpu_baseline(deck, dev, settled_state, protocol="bounded", run_s=3.0):
    """The CPU restatement (OpenMP, all host threads) on the settled state,
    compiled for float and for double.  protocol "full" follows the plan:
    100 warm-up steps, then 3 runs of 200 timed steps, median.  "bounded"
    (the default, a few seconds): 3 warm-up steps, then 3 runs of as many
    steps as take about run_s seconds, median.  The step loop alone is timed
    (steady_clock inside the restatement).  value is the FP32 rate."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    from oracle_py import Oracle
    d = dict(deck)
    d["pos"] = settled_state["pos"]
    d["vel"] = settled_state["vel"]
    d["avel"] = settled_state["avel"]
    d["ptype"] = settled_state["ptype"]
    d["pid"] = settled_state["pid"]
    # radius and mass travel with the particle: the deck's arrays are indexed
    # by the original slot, which is the P_ID (dead particles keep slot 0's)
    slot = np.clip(settled_state["pid"], 0, None)
    d["radius"] = deck["radius"][slot]
    d["mass"] = deck["mass"][slot]
    n = deck["sim"].Num_Particles
    out = {}
    for bits in (32, 64):
        ora = Oracle(d, bits=bits)
        ora.set_active_world_objects(len(deck["wobj"]) - 1)  # hatch open
        if protocol == "full":
            ora.time_steps(100)
            steps = 200
        else:
            t1 = ora.time_steps(3) / 3.0
            steps = max(1, min(200, int(run_s / max(t1, 1e-4))))
        runs = [ora.time_steps(steps) for _ in range(3)]
        ora.close()
        el = float(np.median(runs))
        out[bits] = dict(rate=n * steps / el, steps=steps, runs_s=[round(x, 4) for x in runs])
        log(f"[bench] CPU FP{bits}: {out[bits]['rate']:.3e} particle-steps/s ({steps} steps x 3, median)")
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    warm = "100 warm-up steps, 3 runs of 200 timed steps" if protocol == "full" else \
        f"3 warm-up steps, 3 runs of {out[32]['steps']} (FP32) / {out[64]['steps']} (FP64) timed steps"
    return dict(value=out[32]["rate"], unit="particle-steps/s", cores=threads, kind="port",
                fp64_value=out[64]["rate"], nproc=os.cpu_count(), omp_num_threads=os.environ.get("OMP_NUM_THREADS"),
                protocol=protocol, runs_s={"fp32": out[32]["runs_s"], "fp64": out[64]["runs_s"]},
                sample=f"{warm}, median, of the settled {n}-sphere silo with the hatch open (tangential springs "
                       f"start at zero); FP32 and FP64 OpenMP restatement on {threads} threads")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--settle", type=int, default=300000,
                    help="untimed settle steps, outlet closed: the most to run (it stops once quiet)")
    ap.add_argument("--quiet-vmax", type=float, default=10.0,
                    help="settle until the fastest particle moves below this speed (cm/s)")
    ap.add_argument("--fill", choices=("random", "fcc"), default="random",
                    help="initial state: seeded random non-overlapping fill (the configuration) or an FCC lattice")
    ap.add_argument("--settle-dt", type=float, default=2e-5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-protocol", choices=("bounded", "full"), default="bounded",
                    help="CPU baseline: bounded (a few seconds per precision) or full (100 warm-up steps, "
                         "3 x 200 timed steps, median)")
    ap.add_argument("--skin", type=float, default=None,
                    help="skin masks of the step's stencil scan in cm (default: off, every candidate tested)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed steps, run the library's plain copy kernel (k_copy_u4, "
                         "256 MiB read + 256 MiB written) three times, to scale HBM byte counters")
    ap.add_argument("--save-state", default=None, help="write the settled state (npz) for profiling runs")
    ap.add_argument("--revs", type=float, default=1.0, help="mill workload: revolutions to run")
    ap.add_argument("--workload", choices=("silo", "poly", "hullmill", "mill", "batch", "adjoint"), default="silo",
                    help="silo: the 1M-sphere headline; poly: 200k convex polyhedra (reference law, box); "
                         "hullmill: 200k per-particle convex hulls in a lifter drum; "
                         "batch: 512 independent boxes advanced together vs one after another")
    ap.add_argument("--instances", type=int, default=512, help="batch workload: number of boxes")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="silo workload: skip the adjoint-vs-forward leg reported beside the headline")
    ap.add_argument("--no-batch", action="store_true",
                    help="silo workload: skip the batched-vs-sequential leg reported beside the headline")
    ap.add_argument("--load-state", default=None,
                    help="start from a state written by --save-state instead of settling "
                         "(tangential springs restart at zero)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "poly":
        return run_poly(args)
    if args.workload == "mill":
        return run_mill(args)
    if args.workload == "hullmill":
        return run_hullmill(args)
    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "adjoint":
        return run_adjoint(args)

    import numpy as np
    import bdem
    from bdem import scenarios as S

    t0 = time.time()
    deck = S.silo_deck(args.n, fill=args.fill)
    sim = deck["sim"]
    dt = sim.InitalDelta_t
    dev = make_device(bdem, deck)
    if args.skin is not None:
        dev.set_skin(args.skin)
    ncells = dev.num_cells()
    log(f"[bench] silo n={args.n} cells={ncells} setup {time.time() - t0:.1f}s")

    if args.load_state:
        z = np.load(args.load_state)
        dev.set_radii(z["radius"], z["mass"])
        dev.set_state(z["pos"], z["ptype"], z["pid"], vel=z["vel"], avel=z["avel"])
        args.settle = int(z["settle"])
        log(f"[bench] loaded settled state {args.load_state}")
    else:
        # settle with the outlet closed at a larger time step
        sim.InitalDelta_t = args.settle_dt
        dev.update_sim_info(sim)
        done = 0
        chunk = 5000
        while done < args.settle:
            m = min(chunk, args.settle - done)
            dev.step(m)
            done += m
            s_ = dev.particle_stats()
            log(f"[bench] settle {done} steps ({time.time() - t0:.1f}s): max |v| {s_['max_speed'] * 100:.1f} cm/s, "
                f"mean {s_['mean_speed'] * 100:.2f} cm/s")
            if s_["max_speed"] * 100.0 < args.quiet_vmax:
                break
        args.settle = done
    st = dev.get_state()
    if args.save_state:
        slot = np.clip(st["pid"], 0, None)
        np.savez(args.save_state, pos=st["pos"], vel=st["vel"], avel=st["avel"], ptype=st["ptype"],
                 pid=st["pid"], radius=deck["radius"][slot], mass=deck["mass"][slot],
                 settle=args.settle)
        log(f"[bench] saved settled state to {args.save_state}")
    live = st["pid"] >= 0
    vmax = float(np.max(np.linalg.norm(st["vel"][live], axis=1)))
    top = float(np.max(st["pos"][live, 1]))
    log(f"[bench] settled: bed top y={top:.1f} cm, max |v| {vmax:.1f} cm/s")

    # open the outlet, back to the configured time step
    sim.InitalDelta_t = dt
    dev.update_sim_info(sim)
    dev.open_hatch()
    _, nn = dev.debug_contacts()
    k = float(nn[live].mean())
    dev.step(args.warmup)
    dev.sync()
    n_live = int(np.count_nonzero(dev.get_state()["pid"] >= 0))
    log(f"[bench] warmup {args.warmup} done; live {n_live}, mean contacts/particle {k:.2f}")

    dev.sync()
    dev.timing_begin(args.steps)
    ta = time.perf_counter()
    dev.step(args.steps)
    dev.sync()
    tb = time.perf_counter()
    stages, nrec = dev.timing_end()
    elapsed = tb - ta
    ms_per_step = elapsed * 1e3 / args.steps
    value = n_live * args.steps / elapsed
    status = dev.status()
    log(f"[bench] timed {args.steps} steps: {ms_per_step:.4f} ms/step; stage ms/step "
        + ", ".join(f"{s}={v / nrec:.4f}" for s, v in stages.items()) + f"; status {status}")

    step_ms = stages["step"] / nrec
    kb = step_kernel_bytes(n_live, k, ncells)
    achieved = kb / (step_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "k_step_silo1m_traffic.txt")
    if os.path.exists(tfile):
        for line in open(tfile):
            if line.startswith("traffic_bytes_per_launch:"):
                traffic = float(line.split(":")[1])
    valu_frac = None
    vfile = os.path.join(ROOT, "profiles", "k_step_silo1m_valu.txt")
    if os.path.exists(vfile):
        for line in open(vfile):
            if line.startswith("valu_frac:"):
                valu_frac = float(line.split(":")[1])
    tot_b = step_total_bytes(n_live, k, ncells)
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                kernel="k_step<HERTZ_MINDLIN> (fused contact+force+wall+integrate)",
                measured_bound="valu issue", valu_frac=valu_frac,
                algorithmic_bytes_per_launch=int(kb), avg_launch_ms=round(step_ms, 5))
    step_roof = dict(bytes_per_step=int(tot_b),
                     achieved=round(tot_b / (ms_per_step * 1e-3) / 1e9, 1),
                     frac=round(tot_b / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))

    if args.copy_calibration:
        dev.calibration_copy(256 * 2**20, 3)
        log("[bench] copy calibration: 3 x k_copy_u4 over 256 MiB")

    # size of the incremental sort's changed list over a few more (untimed) steps
    chg = []
    for _ in range(8):
        dev.step(1)
        chg.append(dev.sort_changed())
    changed = float(np.mean(chg))
    log(f"[bench] changed list per step (untimed sample): mean {changed:.0f}, max {max(chg)}")

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (FP32 restatement)...")
        cpu = cpu_baseline(deck, dev, dev.get_state(), protocol=args.cpu_protocol)
        log(f"[bench] CPU baseline {cpu['value']:.3e} (FP32) / {cpu['fp64_value']:.3e} (FP64) particle-steps/s "
            f"on {cpu['cores']} threads")
    dev.close()

    batched = None
    if not args.no_batch:
        batched = batch_leg(count=args.instances)
    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_leg()

    out = {
        "metric": "particle-steps/s at 1M spheres (plain step, silo discharge)",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "silo discharge, 1,048,576 polydisperse spheres, Hertz-Mindlin + "
                               "tangential history, dt 5e-6 s, outlet open",
                   "particles": n_live, "cells": ncells, "mean_contacts": round(k, 3),
                   "settle_steps": args.settle, "settle_dt": args.settle_dt,
                   "stage_ms_per_step": {s: round(v / nrec, 5) for s, v in stages.items()},
                   "changed_per_step": round(changed)},
        "roofline": roof,
        "step_roofline": step_roof,
        "cpu_baseline": cpu,
        "batched": batched,
        "adjoint": adjoint,
    }
    print(json.dumps(out), flush=True)


def batch_leg(count=512, steps=200, warmup=4, tune_steps=5, log_every=64):
    """Batched independent boxes (512 by default: counts log-uniform in
    [1024, 65536], r U[0.2, 0.6] cm, packing 0.3, per-box kn, e, mu; LSD +
    Coulomb + history; dt 1e-5 s) advanced together, against the same boxes
    run one after another in plain contexts.  The autotuner picks the launch
    shape and cell factor first; both sides then use the chosen cells.
    Returns the batched rate, the sequential rate and their ratio."""
    import numpy as np
    import bdem
    from bdem import scenarios as S

    t0 = time.time()
    decks = S.batch_box_decks(count=count, seed=4)
    ntot = sum(d["sim"].Num_Particles for d in decks)
    log(f"[batch] {count} boxes, {ntot} spheres generated in {time.time() - t0:.1f}s")
    batch = bdem.Batch()
    for d in decks:
        dev = make_device(bdem, d)
        batch.add(dev)
        dev.close()
    batch.finalize()
    best, tuned_ms = batch.autotune(steps=tune_steps)
    log(f"[batch] autotuner: {best} ({tuned_ms:.3f} ms/step in its trial)")
    batch.step(warmup)
    batch.sync()
    ta = time.perf_counter()
    batch.step(steps)
    batch.sync()
    t_batch = time.perf_counter() - ta
    status = batch.status()
    log(f"[batch] batched: {steps} steps in {t_batch:.3f} s ({t_batch * 1e3 / steps:.3f} ms/step), status {status}")
    sims = [batch.instance_sim(k) for k in range(count)]
    batch.close()
    t_seq = 0.0
    for k, d in enumerate(decks):
        dd = dict(d)
        dd["sim"] = sims[k]
        dev = make_device(bdem, dd)
        dev.step(warmup)
        dev.sync()
        ta = time.perf_counter()
        dev.step(steps)
        dev.sync()
        t_seq += time.perf_counter() - ta
        dev.close()
        if (k + 1) % log_every == 0:
            log(f"[batch] sequential plain runs: {k + 1}/{count} done ({t_seq:.2f} s of stepping)")
    rate_b = ntot * steps / t_batch
    rate_s = ntot * steps / t_seq
    log(f"[batch] sequential: {t_seq:.3f} s; speedup {t_seq / t_batch:.2f}x")
    return dict(workload=f"{count} independent box fills, {ntot} spheres (counts log-uniform in [1024, 65536]), "
                         "LSD + Coulomb + history, per-box kn/e/mu, dt 1e-5 s",
                instances=count, particles=ntot, steps=steps, tuning=best,
                batched_particle_steps_per_s=round(rate_b, 1), batched_ms_per_step=round(t_batch * 1e3 / steps, 5),
                sequential_particle_steps_per_s=round(rate_s, 1),
                sequential_ms_per_step=round(t_seq * 1e3 / steps, 5),
                speedup_vs_sequential=round(t_seq / t_batch, 3))


def adjoint_leg(n=20000, nsteps=2000, window=100, every=100, settle=10000, iters=1):
    """Differentiable drum calibration: n spheres (radii U[0.15, 0.25] cm)
    in a smooth drum (R 15 cm, L 10 cm, 30 rpm), reference pair law, settled
    (untimed), then nsteps steps at dt 1e-5 s with checkpoints every `every`
    steps.  The loss is the MSE of the windowed bed centre of mass and
    inclination against a target run with hidden parameters; `iters`
    evaluations of loss and gradient (with normalised steps between them).
    Reports the forward and adjoint times against plain steps.  Over the
    full horizon the exact gradient of a resting bed is huge: the reference
    friction law's tangential force has a velocity Jacobian of order
    mu (m/dt) / |v_rel|, which amplifies the cotangent at every step."""
    import numpy as np
    import bdem
    from bdem import adjoint as ADJ
    from bdem import scenarios as S

    t0 = time.time()
    deck = S.drum_deck(n=n, seed=5, R=15.0, L=10.0, rpm=30.0, kn=1e5, e=0.5, mu=0.4, roll_res=0.002)
    dev = make_device(bdem, deck)
    dev.drum_rotation(True)
    dev.step(settle)
    st = dev.get_state()
    order = np.argsort(st["pid"])
    dev.close()
    d = dict(deck)
    d["pos"], d["vel"], d["avel"] = st["pos"][order], st["vel"][order], st["avel"][order]
    log(f"[adjoint] drum n={n} settled {settle} steps ({time.time() - t0:.1f}s)")
    # plain steps on the same state and parameters, for the cost ratio
    plain = make_device(bdem, d)
    plain.drum_rotation(True)
    plain.step(2)
    plain.sync()
    ta = time.perf_counter()
    plain.step(nsteps)
    plain.sync()
    plain_ms = (time.perf_counter() - ta) * 1e3
    plain.close()
    dev = make_device(bdem, d)
    dev.drum_rotation(True)
    run = ADJ.AdjointRun(dev, nsteps, window, checkpoint_every=every)
    hidden = np.array([0.55, 0.35, 1.6e5, 0.01])
    theta = np.array([0.4, 0.5, 1e5, 0.002])
    target = run.forward(hidden)
    hist = []
    fwd_ms = bwd_ms = None
    for it in range(iters):
        obs = run.forward(theta)
        r = obs - target
        loss = float(np.mean(r * r))
        g = run.backward(2.0 * r / len(r))
        fwd_ms, bwd_ms = run.timing()
        hist.append(dict(iter=it, loss=loss, theta=[float(x) for x in theta], grad=[float(x) for x in g]))
        log(f"[adjoint] iter {it}: loss {loss:.4e} theta {theta} grad {g} (forward {fwd_ms:.1f} ms, "
            f"adjoint {bwd_ms:.1f} ms)")
        # normalised gradient step: each parameter moves by at most 5 % of itself
        step = -0.05 * theta * np.sign(g) * np.minimum(1.0, np.abs(g * theta) / max(np.max(np.abs(g * theta)), 1e-300))
        theta = np.clip(theta + step, [0.05, 0.31, 1e4, 0.0], [1.0, 0.89, 1e6, 0.2])
    run.close()
    dev.close()
    return dict(workload=f"drum calibration: {n} spheres r U[0.15, 0.25] cm, R 15 cm, L 10 cm, 30 rpm, "
                         f"{nsteps} steps at dt 1e-5 s, checkpoints every {every}, windows of {window}",
                plain_ms=round(plain_ms, 3), forward_ms=round(fwd_ms, 3), adjoint_ms=round(bwd_ms, 3),
                adjoint_over_forward=round(bwd_ms / fwd_ms, 3), adjoint_over_plain=round(bwd_ms / plain_ms, 3),
                hidden_theta=[float(x) for x in hidden], calibration=hist)


def run_adjoint(args):
    a = adjoint_leg()
    out = {
        "metric": "adjoint step cost vs forward, differentiable drum calibration",
        "value": a["adjoint_over_forward"],
        "unit": "x forward",
        "n_gpus": 1,
        "steps": 2000,
        "warmup": 0,
        "ms_per_step": round(a["adjoint_ms"] / 2000, 5),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": a["workload"], "adjoint": a},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_batch(args):
    steps = args.steps if args.steps != 1000 else 200
    b = batch_leg(count=args.instances, steps=steps)
    out = {
        "metric": "particle-steps/s, batched independent simulations (vs sequential plain runs)",
        "value": b["batched_particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 4,
        "ms_per_step": b["batched_ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": b["workload"], "batched": b},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_poly(args):
    """Polyhedra throughput: 200,000 seeded convex hulls of 12 points
    (20 triangular faces, 30 edges, circumradius ~0.48 cm) on a jittered
    lattice in a closed box, random orientations, reference pair law; the
    four polyhedra kernels (rotate, list + queue, pairs, finish) plus the
    sort, timed per step."""
    import numpy as np
    import bdem
    from bdem import scenarios as S

    n = args.n if args.n != 1048576 else 200000
    t0 = time.time()
    side = int(np.ceil(n ** (1.0 / 3.0)))
    spacing = 0.85
    L = side * spacing + 2.5
    deck = S.poly_deck(n, seed=71, L=L, shape="hull", size=1.0, spacing=spacing, vmax=10.0, wmax=5.0)
    dev = bdem.Device(deck["wobj"], deck["pobj"], deck["sim"])
    dev.set_state(deck["pos"], deck["ptype"], deck["pid"], vel=deck["vel"], avel=deck["avel"],
                  quat=deck["quat"])
    ncells = dev.num_cells()
    _, nn = dev.debug_contacts()
    k = float(nn.mean())
    log(f"[bench] poly n={n} box {L:.1f} cm cells={ncells} candidates/particle {k:.2f} "
        f"setup {time.time() - t0:.1f}s")
    dev.step(args.warmup)
    dev.sync()
    dev.timing_begin(args.steps)
    ta = time.perf_counter()
    dev.step(args.steps)
    dev.sync()
    tb = time.perf_counter()
    stages, nrec = dev.timing_end()
    elapsed = tb - ta
    ms_per_step = elapsed * 1e3 / args.steps
    st = dev.get_state()
    n_live = int(np.count_nonzero(st["pid"] >= 0))
    finite = bool(np.all(np.isfinite(st["pos"])))
    status = dev.status()
    log(f"[bench] timed {args.steps} steps: {ms_per_step:.4f} ms/step; stage ms/step "
        + ", ".join(f"{s_}={v / nrec:.4f}" for s_, v in stages.items())
        + f"; status {status}; finite {finite}")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle_py import Oracle
        ora = Oracle(deck, bits=32)
        ora.time_steps(1)
        steps = 3
        el = ora.time_steps(steps)
        ora.close()
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = dict(value=n * steps / el, unit="particle-steps/s", cores=cores, kind="port",
                   sample=f"{steps} steps (after 1 untimed) of the same {n}-hull deck from its initial "
                          f"state, FP32 OpenMP restatement")
        log(f"[bench] CPU baseline {cpu['value']:.3e} particle-steps/s on {cores} threads")
    dev.close()
    out = {
        "metric": "particle-steps/s, convex polyhedra (reference law)",
        "value": round(n_live * args.steps / elapsed, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{n} convex hulls of 12 seeded points (20 faces, 30 edges), jittered "
                               f"lattice {spacing} cm in a closed {L:.1f} cm box, dt 1e-5 s",
                   "particles": n_live, "cells": ncells, "candidates_per_particle": round(k, 3),
                   "stage_ms_per_step": {s_: round(v / nrec, 5) for s_, v in stages.items()}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_hullmill(args):
    """Polyhedral ball mill: 200,000 per-particle convex hulls (8 to 16
    seeded points on a sphere, bounding radius U[0.3, 0.6] cm, 2700 kg/m^3),
    each its own shape-table row, in a 30 cm-radius, 20 cm-long drum with 8
    lifters turning at 60 % of the critical speed; reference pair law kn
    2e5, e 0.4, mu 0.6, dt 2e-6 s.  The hulls start in contact (random fill
    of half-radius spheres) and spread for --settle untimed steps with the
    drum turning, then --steps steps are timed."""
    import numpy as np
    import bdem
    from bdem import scenarios as S

    n = args.n if args.n != 1048576 else 200000
    if args.steps == 5000:        # the silo's default: this configuration runs 20,000 steps
        args.steps = 20000
    if args.settle == 300000:     # the silo's default: settle this charge for 20,000 steps
        args.settle = 20000
    if args.settle_dt == 2e-5:    # the silo's default: hulls this small need the configuration's dt
        args.settle_dt = 2e-6
    t0 = time.time()
    deck = S.hull_mill_deck(n)
    dev = bdem.Device(deck["wobj"], deck["pobj"], deck["sim"], dobj=deck["dobj"])
    dev.set_poly_shapes(deck["shapes"])
    com = np.array([[s_.COM.x, s_.COM.y, s_.COM.z] for s_ in deck["shapes"]], np.float32)
    if args.load_state:   # a state written by --save-state (positions as returned: centre - COM)
        z = np.load(args.load_state)
        dev.set_state(z["pos"] + com[np.maximum(z["pid"], 0)], z["ptype"], z["pid"], vel=z["vel"], avel=z["avel"],
                      quat=z["quat"])
        log(f"[bench] loaded state {args.load_state}")
    else:
        dev.set_state(deck["pos"], deck["ptype"], deck["pid"], vel=deck["vel"], avel=deck["avel"],
                      quat=deck["quat"])
    dev.drum_rotation(True)
    ncells = dev.num_cells()
    log(f"[bench] hullmill n={n} cells={ncells} rpm {deck['rpm']:.2f} setup {time.time() - t0:.1f}s")
    dt = deck["sim"].InitalDelta_t
    if args.settle > 0:
        sim = deck["sim"]
        sim.InitalDelta_t = float(args.settle_dt)
        dev.update_sim_info(sim)
        dev.set_lifter_rotation(deck["rad_per_step"] * args.settle_dt / dt, deck["lifter_about"])
        done = 0
        while done < args.settle:
            k = min(5000, args.settle - done)
            dev.step(k)
            done += k
            st = dev.particle_stats()
            log(f"[bench] settle {done} steps at dt {args.settle_dt:g}: mean speed {st['mean_speed']:.3f} m/s, "
                f"max {st['max_speed']:.3f} m/s, status {dev.status()}")
        sim.InitalDelta_t = float(dt)
        dev.update_sim_info(sim)
    if args.save_state:
        st = dev.get_state()
        np.savez(args.save_state, pos=st["pos"], vel=st["vel"], avel=st["avel"], quat=st["quat"], ptype=st["ptype"],
                 pid=st["pid"])
        log(f"[bench] saved state to {args.save_state}")
    dev.set_lifter_rotation(deck["rad_per_step"], deck["lifter_about"])
    _, nn = dev.debug_contacts()
    k = float(nn.mean())
    dev.step(args.warmup)
    dev.sync()
    dev.timing_begin(args.steps)
    ta = time.perf_counter()
    done = 0
    while done < args.steps:
        kk = min(5000, args.steps - done)
        dev.step(kk)
        done += kk
        if done < args.steps:
            dev.sync()
            log(f"[bench] timed {done} of {args.steps} steps")
    dev.sync()
    tb = time.perf_counter()
    stages, nrec = dev.timing_end()
    elapsed = tb - ta
    ms_per_step = elapsed * 1e3 / args.steps
    st = dev.get_state()
    n_live = int(np.count_nonzero(st["pid"] >= 0))
    finite = bool(np.all(np.isfinite(st["pos"])))
    status = dev.status()
    queued, narrow = dev.debug_poly_queues()
    log(f"[bench] timed {args.steps} steps: {ms_per_step:.4f} ms/step, candidates/particle {k:.2f}; stage "
        "ms/step " + ", ".join(f"{s_}={v / nrec:.4f}" for s_, v in stages.items())
        + f"; status {status}; finite {finite}; last step: {queued} listed pairs, {narrow} to the wave-per-pair kernel")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle_py import Oracle
        com = np.array([[s_.COM.x, s_.COM.y, s_.COM.z] for s_ in deck["shapes"]], np.float32)
        live = np.maximum(st["pid"], 0)
        state = dict(deck, pos=st["pos"] + com[live], vel=st["vel"], avel=st["avel"], quat=st["quat"],
                     ptype=st["ptype"], pid=st["pid"])
        ora = Oracle(state, bits=32)
        ora.set_drum_rotation(True)
        ora.run(1, deck["rad_per_step"], deck["lifter_about"])
        steps = 3
        tc = time.perf_counter()
        ora.run(steps, deck["rad_per_step"], deck["lifter_about"])
        el = time.perf_counter() - tc
        ora.close()
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = dict(value=n * steps / el, unit="particle-steps/s", cores=cores, kind="port",
                   sample=f"{steps} steps (after 1 untimed) from the state the timed run ended in, lifters "
                          f"turning; FP32 OpenMP restatement")
        log(f"[bench] CPU baseline {cpu['value']:.3e} particle-steps/s on {cores} threads")
    dev.close()
    out = {
        "metric": "particle-steps/s, per-particle convex hulls in a lifter drum (reference law)",
        "value": round(n_live * args.steps / elapsed, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{n} per-particle convex hulls of 8-16 seeded points (bounding radius "
                               "U[0.3, 0.6] cm), drum R 30 cm L 20 cm, 8 lifters 2 x 3 cm, "
                               f"{deck['rpm']:.2f} rpm (60% critical), kn 2e5, e 0.4, mu 0.6, dt 2e-6 s",
                   "settle": {"steps": args.settle, "dt": args.settle_dt},
                   "particles": n_live, "cells": ncells, "candidates_per_particle": round(k, 3),
                   "last_step_pairs": {"listed": queued, "wave_per_pair": narrow},
                   "stage_ms_per_step": {s_: round(v / nrec, 5) for s_, v in stages.items()}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_mill(args):
    """The reference's ball-mill deck (Projects/3.BallMill, World Mill_SPH30:
    242 steel balls of 5 cm in a 90 cm drum with 8 lifters, 22 RPM, dt 1e-5 s)
    from its start file, --revs revolutions (272,727 steps each) with the lifters
    turned on the device and the dissipation tallies on; the rate is
    particle-steps/s and the mill power after the revolution is reported
    beside it (after one revolution: the reference's known answer)."""
    import numpy as np
    from bdem.deck import Deck

    d = Deck(os.path.join(ROOT, "tests", "golden", "decks", "3.BallMill"), "Mill_SPH30")
    dev = d.device()
    d.load_state(dev)
    dev.reset_energy()
    dev.drum_rotation(True)
    a = d.info.rotate_about
    dev.set_lifter_rotation(d.info.rad_per_step, (a.x, a.y, a.z))
    steps = max(1, int(round(args.revs * d.info.steps_per_rev)))
    n = d.sim.Num_Particles
    dev.sync()
    ta = time.perf_counter()
    dev.step(steps)
    dev.sync()
    tb = time.perf_counter()
    elapsed = tb - ta
    power = dev.energy().astype(np.float64) / (steps * d.sim.InitalDelta_t)
    log(f"[bench] mill {steps} steps in {elapsed:.2f} s; power particle {power[0]:.2f} lifter {power[1]:.2f} "
        f"drum {power[2]:.2f} total {power.sum():.2f} W; status {dev.status()}")
    dev.close()
    out = {
        "metric": "particle-steps/s, ball-mill deck Mill_SPH30",
        "value": round(n * steps / elapsed, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": round(elapsed * 1e3 / steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "reference project deck",
        "config": {"workload": "Mill_SPH30: 242 balls of 5 cm, 90 cm drum, 8 lifters, 22 RPM, dt 1e-5 s, "
                               "from StartMill_SPH30, lifters turned on the device, tallies on",
                   "revolutions": round(steps / d.info.steps_per_rev, 4),
                   "power_W": {"particle": round(power[0], 3), "lifter": round(power[1], 3),
                               "drum": round(power[2], 3), "total": round(power.sum(), 3)},
                   "published_power_W": {"particle": 319.724, "lifter": 156.378, "drum": 149.878,
                                         "total": 625.98}},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
