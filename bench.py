#!/usr/bin/env python3
"""Benchmark: cloth substeps/s at a fixed residual, warm-start + batched vs
plain, on one MI355X.

Workload: the 512-cloth batch -- grid side U{16..256} vertices, aspect
U(0.5, 2), 1 m wide, 0-4 random boundary pins, its own sphere collider
(r U(0.2, 0.4) m) each -- packed into one context; dt = 1/600 s (10 substeps
per 1/60 s frame), distance constraints, colour-ordered Gauss-Seidel in LDS
tiles.  One "step" is one substep of the whole batch.  Inputs are generated
on the host from a seeded RNG and are resident in HBM before timing starts.

Protocol (every timed run starts from the same device snapshot):
  1. learned warm start on the 32x32 scene (wind U(0,5) m/s, random pins and
     sphere): trained on ~200k solver-generated samples, then evaluated on
     held-out scenes as the sweeps needed to reach each cloth's plain
     residual level with and without it (config.warm_start_32x32);
  2. the batch: the autotuner picks block size, sweeps per launch and tile
     size (results do not depend on the choice); the warm-start network for
     the batch is trained on samples the solver generates during the
     batch's own warm-up substeps, before the snapshot;
  3. plain solver: 20 sweeps per substep (the reference's fixed budget,
     SolverPBD::ProjectConstraint(st, iterations)); its mean end-of-substep
     residual (max |d-L|/L per cloth, averaged over cloths and substeps) is
     the fixed residual R.  Batched and unbatched plain runs give the same
     positions bit for bit, hence the same R;
  4. warm-started solver: a cloth whose prediction is kept runs a short
     budget k, a cloth whose prediction is discarded runs the plain 20
     sweeps; k is the smallest budget whose mean residual over the same
     substeps is <= R (bisection);
  5. timed: the plain solver unbatched (each cloth stepped alone, as the
     reference's per-object SolverPBD), the plain solver batched, and the
     warm-started batched solver, both batched runs replayed from captured
     graphs;
  6. the colour-free Jacobi variant as a baseline (config.jacobi_baseline):
     the smallest sweep count reaching R, and its substeps/s.
`value` is the warm-started batched solver's substeps/s;
config.speedup_vs_plain divides it by the unbatched plain solver's rate and
config.speedup_vs_plain_batched by the batched plain solver's.

`roofline` is for the dominant kernel (the LDS-tile Gauss-Seidel kernel),
from a second, eager pass over the same warm-started substeps with HIP events
around every launch on the library's own stream: algorithmic bytes (68 B per
distance prim per sweep actually run) over the summed launch time.
`cpu_baseline` times the CPU oracle (one thread, plain solver, 20 sweeps) on
a bounded sample of the same batch, scaled to the whole batch by
vertex-sweeps.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
PLAIN_SWEEPS = 20


def traffic_key(cfg):
    """Launch configuration key of the committed tile-kernel traffic figures."""
    key = "ts{tile_sweeps}_bt{block_threads}_mv{tile_max_verts}".format(**cfg)
    return key + ("_sw1" if cfg.get("split_whole") else "")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed substeps per solver")
    ap.add_argument("--warmup", type=int, default=12, help="plain substeps before the training samples")
    ap.add_argument("--match-substeps", type=int, default=10, help="substeps of the residual comparison")
    ap.add_argument("--tile-sweeps", type=int, default=4, help="sweeps per launch when not autotuned")
    ap.add_argument("--count", type=int, default=512, help="cloths in the batch")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--train-scenes", type=int, default=24, help="32x32 scenes for the reference-scene network")
    ap.add_argument("--eval-scenes", type=int, default=16, help="held-out 32x32 scenes")
    ap.add_argument("--collect-substeps", type=int, default=2, help="batch substeps recorded as training samples")
    ap.add_argument("--target-sweeps", type=int, default=60, help="sweeps behind each training target")
    ap.add_argument("--batch-epochs", type=int, default=4)
    ap.add_argument("--scene-epochs", type=int, default=200, help="epochs of the 32x32-scene network")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--unbatched-steps", type=int, default=3, help="timed substeps of the unbatched plain solver")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (use --tile-sweeps, 1024 threads)")
    ap.add_argument("--save-network", metavar="NPZ", help="write the batch network and the settings that "
                    "positionbaseddynamicsold_amd.tilework needs to repeat the profiled passes")
    ap.add_argument("--warm-rule", type=int, default=1, choices=(0, 1),
                    help="fallback rule of the headline run: 0 max |d-L|/L, 1 RMS (the other is reported too)")
    ap.add_argument("--jacobi-omega", type=float, default=1.5,
                    help="relaxation of the Jacobi baseline (constraint averaging)")
    ap.add_argument("--jacobi-max-sweeps", type=int, default=400, help="largest Jacobi sweep count tried")
    ap.add_argument("--no-jacobi", action="store_true", help="skip the Jacobi baseline")
    ap.add_argument("--warm-gate", type=float, default=1e-4,
                    help="no correction is predicted for a vertex whose one-ring prims are all within this "
                         "relative strain (0: every vertex); the run without the gate is reported too")
    return ap.parse_args()


def scene_network(capi, scenes, n_train, n_eval, target_sweeps, epochs):
    """The 32x32 reference scene: ~200k samples from n_train seeds, then the
    mean sweeps to each held-out cloth's plain residual level (its mean
    20-sweep residual) with and without the warm start."""
    t = time.perf_counter()
    ctx = capi.Context(0)
    for s in range(n_train):
        scenes.add_to_context(ctx, scenes.config4(s))
    ctx.init()
    ctx.set_solver(dt=scenes.DT, iterations=PLAIN_SWEEPS)
    ctx.ws_init(0)
    for _ in range(8):
        ctx.step(30)
        ctx.ws_collect(target_sweeps, 1)
    n = ctx.ws_num_samples()
    loss = ctx.ws_train(epochs=epochs, batch=2048, lr=1e-3, seed=1)
    prm, norm = ctx.ws_params(), ctx.ws_norm()
    ctx.close()
    ev = capi.Context(0)
    ids = [scenes.add_to_context(ev, scenes.config4(1000 + s)) for s in range(n_eval)]
    ev.init()
    ev.set_solver(dt=scenes.DT, iterations=PLAIN_SWEEPS)
    ev.ws_set_params(prm)
    ev.ws_set_norm(norm)
    ev.step(120)
    st = [(ev.get(i, 0), ev.get(i, 1)) for i in ids]
    tau = np.zeros(len(ids))
    for _ in range(10):
        ev.step(1)
        tau += [ev.residual(i)[0] for i in ids]
    tau /= 10
    mean = {}
    for ws in (False, True):
        for i, (x, v), tv in zip(ids, st, tau):
            ev.set(i, 0, x)
            ev.set(i, 1, v)
            ev.set_tolerance(i, float(tv))
        ev.ws_enable(ws)
        ev.set_solver(dt=scenes.DT, iterations=80, tol=0.0, check_every=1)
        tot = 0
        for _ in range(40):
            ev.step(1)
            tot += sum(ev.last_sweeps(i) for i in ids)
        mean[ws] = tot / (40 * len(ids))
    ev.close()
    info = {"samples": n, "epochs": epochs, "target_sweeps": target_sweeps, "standardised_loss": loss,
            "zero_predictor_loss": 1.5,
            "held_out_scenes": n_eval, "mean_sweeps_plain": mean[False], "mean_sweeps_warm": mean[True],
            "seconds": time.perf_counter() - t,
            "definition": "mean sweeps per cloth-substep to reach the cloth's own mean plain 20-sweep residual, "
                          "residual checked after every sweep, 40 substeps from the same state"}
    log(f"[bench] 32x32 scene: network from {n} samples (loss {loss:.3f}); sweeps to the plain residual level: "
        f"plain {mean[False]:.2f}, warm {mean[True]:.2f}")
    return info


def cpu_baseline(specs, sweeps, budget_s):
    """Times the oracle (1 thread) on the first cloths of the batch until the
    budget is spent; returns vertex-sweeps/s and a description."""
    from oracle import oracle as orc

    sample, nv = [], 0
    for s in specs[:16]:
        pos, tris = orc.grid_mesh(s.rx, s.ry, s.sx, s.sy, jitter=s.jitter, seed=s.seed)
        oc = orc.OracleCloth(pos, tris, orc.grid_pairs(s.rx, s.ry), pins=s.pins, iterations=sweeps)
        prm = oc.params(dt=1.0 / 600.0, sphere=s.sphere)
        sample.append((oc, prm))
        nv += s.nv
        if nv > 60000:
            break
    for oc, prm in sample:
        oc.substep(prm, cap=sweeps)
    done_vs, n_sub, t0 = 0, 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        for oc, prm in sample:
            oc.substep(prm, cap=sweeps)
            done_vs += oc.nv * sweeps
        n_sub += 1
    el = time.perf_counter() - t0
    desc = (f"CPU oracle (plain solver, {sweeps} sweeps), 1 thread, first {len(sample)} cloths of the batch "
            f"({nv} vertices), {n_sub} substeps each in {el:.1f} s; scaled to the full batch by vertex-sweeps")
    return done_vs / el, desc


def jacobi_baseline(a, capi, scenes, ctx, ids, r_plain, warm_rate, plain_solver):
    """The colour-free Jacobi variant (constraint averaging) on the same
    constraint problems as the plain solver: every substep of the plain
    trajectory from the snapshot is solved again from its start state (x, v)
    with S Jacobi sweeps; S is the smallest count whose mean end-of-substep
    residual over those problems is <= R (bisection).  Its substeps/s are
    timed from a captured graph.  (Along its own trajectory the max-strain
    residual does not compare solver types: a solver that barely propagates
    the pins' support lets the cloth free-fall with little strain.)"""
    ctx.snapshot(restore=True)
    plain_solver()
    starts = []
    for _ in range(a.match_substeps):
        starts.append(ctx.state_save())
        ctx.step(1)

    def solver(sweeps):
        ctx.ws_enable(False)
        ctx.set_solver(dt=scenes.DT, iterations=sweeps, path=capi.PATH_JACOBI, jacobi_omega=a.jacobi_omega,
                       graph=1)

    tried = {}

    def mean_residual(sweeps):
        solver(sweeps)
        acc = 0.0
        for x, v in starts:
            ctx.state_load(x, v)
            ctx.step(1)
            acc += float(ctx.residual_all()[0].mean())
        tried[sweeps] = acc / len(starts)
        return tried[sweeps]

    lo, hi = 1, a.jacobi_max_sweeps
    reached = mean_residual(hi) <= r_plain
    if reached:
        while hi > lo:
            mid = (lo + hi) // 2
            if mean_residual(mid) <= r_plain:
                hi = mid
            else:
                lo = mid + 1
        mean_residual(hi)
    del starts
    ctx.snapshot(restore=True)
    solver(hi)
    ctx.step(1)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(a.steps)
    ctx.sync()
    el = time.perf_counter() - t0
    infos = [ctx.info(i) for i in ids]
    nv = sum(i.nv for i in infos)
    dist_inc, bend_inc = sum(2 * i.ne for i in infos), sum(4 * i.nb for i in infos)
    out = {"omega": a.jacobi_omega, "sweeps": hi, "reaches_R": reached, "residual": tried[hi],
           "substeps_per_s": a.steps / el, "ms_per_step": el * 1e3 / a.steps,
           "warm_start_speedup_over_jacobi": warm_rate / (a.steps / el),
           "sweeps_tried": {str(k): v for k, v in sorted(tried.items())},
           # k_jacobi: p in/out, cloth id, degree and two offsets per vertex; 8 B per distance
           # incidence, 24 B per bending incidence (neighbour positions are re-reads)
           "alg_bytes_per_sweep": 48 * nv + 8 * dist_inc + 24 * bend_inc,
           "definition": "PBD_PATH_JACOBI: every vertex moves by omega/n times the sum of the corrections its n "
                         "prims give it on the sweep's input positions.  Sweeps: smallest count whose mean "
                         "end-of-substep residual, over the plain trajectory's substeps each solved from the "
                         "same start state, is <= R (stiffness schedule as the plain solver)"}
    log(f"[bench] Jacobi (omega {a.jacobi_omega}): {hi} sweeps {'reach' if reached else 'do not reach'} R "
        f"({tried[hi]:.5e} vs {r_plain:.5e}); {out['substeps_per_s']:.2f} substeps/s ({out['ms_per_step']:.2f} ms); "
        f"tried {out['sweeps_tried']}")
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); sharding this path across GPUs "
              f"is a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)
    from positionbaseddynamicsold_amd import capi, scenes

    scene_info = scene_network(capi, scenes, a.train_scenes, a.eval_scenes, a.target_sweeps, a.scene_epochs)

    specs = scenes.config5(seed=a.seed, count=a.count)
    nv_total = sum(s.nv for s in specs)
    ne_total = sum(s.ne for s in specs)
    log(f"[bench] building {len(specs)} cloths, {nv_total} vertices, {ne_total} distance prims")
    ctx = capi.Context(0)
    ids = [scenes.add_to_context(ctx, s) for s in specs]
    t = time.perf_counter()
    ctx.init()
    log(f"[bench] init (constraints, GPU colouring, tile plan) {time.perf_counter() - t:.1f} s")
    tuned = {"tile_sweeps": a.tile_sweeps, "block_threads": 1024, "tile_max_verts": 0, "split_whole": 0}
    ctx.set_solver(dt=scenes.DT, iterations=PLAIN_SWEEPS, check_every=a.tile_sweeps, tile_sweeps=a.tile_sweeps,
                   graph=1)
    ctx.step(2)
    if not a.no_tune:
        t = time.perf_counter()
        best, _, best_ms, trials = ctx.autotune(2)
        tuned = {"tile_sweeps": best.tile_sweeps, "block_threads": best.block_threads,
                 "tile_max_verts": best.tile_max_verts, "split_whole": best.split_whole, "trials": trials,
                 "best_ms_per_substep": best_ms, "seconds": time.perf_counter() - t}
        log(f"[bench] autotune: {tuned}")

    rule = {"rule": a.warm_rule, "gate": a.warm_gate}
    # launch configuration per run: the plain solver's from the tuning above,
    # the warm-started solver's from a second tuning once its budget is known
    # (results do not depend on it)
    tuned_warm = dict(tuned)

    def solver(sweeps, ws, graph=1, warm=0):
        cfg = tuned_warm if ws else tuned
        ctx.ws_enable(ws)
        ctx.set_solver(dt=scenes.DT, iterations=sweeps, check_every=cfg["tile_sweeps"],
                       tile_sweeps=cfg["tile_sweeps"], tile_max_verts=cfg["tile_max_verts"],
                       block_threads=cfg["block_threads"], graph=graph, warm_iterations=warm,
                       warm_rule=rule["rule"], warm_gate=rule["gate"], split_whole=cfg["split_whole"])

    # warm-up, with the batch's own training samples recorded along the way
    solver(PLAIN_SWEEPS, False)
    ctx.step(a.warmup)
    t = time.perf_counter()
    ctx.ws_init(0)
    for _ in range(a.collect_substeps):
        ctx.ws_collect(a.target_sweeps, 1)
        ctx.step(1)
    n_samples = ctx.ws_num_samples()
    loss = ctx.ws_train(epochs=a.batch_epochs, batch=4096, lr=1e-3, seed=1)
    batch_train = {"samples": n_samples, "epochs": a.batch_epochs, "target_sweeps": a.target_sweeps,
                   "standardised_loss": loss, "zero_predictor_loss": 1.5, "seconds": time.perf_counter() - t,
                   "source": f"the batch's own substeps after {a.warmup} warm-up substeps (before the snapshot): "
                             f"features at the predicted state, target = the correction of {a.target_sweeps} "
                             f"Gauss-Seidel sweeps"}
    log(f"[bench] batch network: {n_samples} samples, {a.batch_epochs} epochs, loss {loss:.3f} "
        f"in {batch_train['seconds']:.1f} s")
    solver(PLAIN_SWEEPS, False)
    ctx.step(2)
    ctx.sync()
    ctx.snapshot()
    log("[bench] snapshot taken")

    def residuals(ws, warm=0):
        ctx.snapshot(restore=True)
        solver(PLAIN_SWEEPS, ws, warm=warm)
        acc, used = np.zeros(len(ids)), 0.0
        for _ in range(a.match_substeps):
            ctx.step(1)
            acc += ctx.residual_all()[0]
            if ws:
                used += sum(ctx.ws_last_used(i) for i in ids) / len(ids)
        return acc / a.match_substeps, used / a.match_substeps

    r_plain_c, _ = residuals(False)
    r_plain = float(r_plain_c.mean())

    def budget(rl, gate):
        """Smallest warm budget k (discarded cloths keep 20) with mean residual <= R."""
        rule["rule"], rule["gate"] = rl, gate
        tried = {}

        def mean_warm(k):
            r, u = residuals(True, warm=k)
            tried[k] = float(r.mean())
            return r, u

        lo, hi = 1, PLAIN_SWEEPS
        while hi > lo:
            mid = (lo + hi) // 2
            if mean_warm(mid)[0].mean() <= r_plain:
                hi = mid
            else:
                lo = mid + 1
        r_c, used = mean_warm(hi)
        ratio = r_c / np.maximum(r_plain_c, 1e-30)
        log(f"[bench] rule {'RMS' if rl else 'max'}, gate {gate:g}: warm start reaches {r_c.mean():.5e} (R = {r_plain:.5e}) "
            f"with budget {hi} (kept) / {PLAIN_SWEEPS} (discarded); tried {dict(sorted(tried.items()))}; kept for "
            f"{100 * used:.0f}% of cloth-substeps; per-cloth residual ratio warm/plain median "
            f"{np.median(ratio):.3f}, p90 {np.quantile(ratio, 0.9):.3f}, max {ratio.max():.3f}")
        return {"rule": "RMS" if rl else "max", "gate": gate, "budget_kept": hi, "budget_discarded": PLAIN_SWEEPS,
                "residual": float(r_c.mean()), "kept_fraction": used,
                "per_cloth_ratio_median": float(np.median(ratio)),
                "per_cloth_ratio_p90": float(np.quantile(ratio, 0.9)),
                "per_cloth_ratio_max": float(ratio.max()),
                "budgets_tried": {str(k): v for k, v in sorted(tried.items())}}

    other = budget(1 - a.warm_rule, a.warm_gate)
    ungated = budget(a.warm_rule, 0.0) if a.warm_gate > 0 else None
    head = budget(a.warm_rule, a.warm_gate)
    warm_k, r_warm, used_frac = head["budget_kept"], head["residual"], head["kept_fraction"]
    if not a.no_tune:
        ctx.snapshot(restore=True)
        solver(PLAIN_SWEEPS, True, warm=warm_k)
        t = time.perf_counter()
        best, _, best_ms, trials = ctx.autotune(2)
        tuned_warm = {"tile_sweeps": best.tile_sweeps, "block_threads": best.block_threads,
                      "tile_max_verts": best.tile_max_verts, "split_whole": best.split_whole, "trials": trials,
                      "best_ms_per_substep": best_ms, "seconds": time.perf_counter() - t}
        log(f"[bench] autotune (warm-started run): {tuned_warm}")
    if a.save_network:
        np.savez(a.save_network, params=ctx.ws_params(), norm=ctx.ws_norm(), count=a.count, seed=a.seed,
                 warmup=a.warmup, collect_substeps=a.collect_substeps, target_sweeps=a.target_sweeps, settle=2,
                 plain_sweeps=PLAIN_SWEEPS, warm_k=warm_k, warm_rule=a.warm_rule, warm_gate=a.warm_gate,
                 tile_sweeps=tuned["tile_sweeps"], tile_max_verts=tuned["tile_max_verts"],
                 block_threads=tuned["block_threads"], split_whole=tuned["split_whole"],
                 warm_tile_sweeps=tuned_warm["tile_sweeps"], warm_tile_max_verts=tuned_warm["tile_max_verts"],
                 warm_block_threads=tuned_warm["block_threads"], warm_split_whole=tuned_warm["split_whole"])

    def timed(ws, steps, warm=0):
        ctx.snapshot(restore=True)
        solver(PLAIN_SWEEPS, ws, warm=warm)
        ctx.step(1)
        ctx.sync()
        ctx.profile_read(reset=True)
        t0 = time.perf_counter()
        ctx.step(steps)
        ctx.sync()
        el = time.perf_counter() - t0
        replays = ctx.profile_read(reset=True).graph_replays
        return steps / el, el * 1e3 / steps, replays

    def timed_unbatched(steps):
        ctx.snapshot(restore=True)
        solver(PLAIN_SWEEPS, False, graph=0)
        for i in ids:
            ctx.step_cloth(i, 1)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            for i in ids:
                ctx.step_cloth(i, 1)
        ctx.sync()
        el = time.perf_counter() - t0
        return steps / el, el * 1e3 / steps

    def profiled(ws, steps, warm=0):
        ctx.snapshot(restore=True)
        solver(PLAIN_SWEEPS, ws, graph=0, warm=warm)
        ctx.step(1)
        ctx.sync()
        ctx.profile(True)
        ctx.profile_read(reset=True)
        ctx.step(steps)
        ctx.sync()
        prof = ctx.profile_read(reset=True)
        ctx.profile(False)
        return prof

    unb_rate, unb_ms = timed_unbatched(a.unbatched_steps)
    log(f"[bench] plain unbatched (one cloth at a time) {unb_rate:.2f} substeps/s ({unb_ms:.2f} ms)")
    plain_rate, plain_ms, plain_replays = timed(False, a.steps)
    rule["rule"] = 1 - a.warm_rule
    other["substeps_per_s"], other["ms_per_step"], _ = timed(True, a.steps, warm=other["budget_kept"])
    rule["rule"] = a.warm_rule
    if ungated is not None:
        rule["gate"] = 0.0
        ungated["substeps_per_s"], ungated["ms_per_step"], _ = timed(True, a.steps, warm=ungated["budget_kept"])
        rule["gate"] = a.warm_gate
    warm_rate, warm_ms, warm_replays = timed(True, a.steps, warm=warm_k)
    log(f"[bench] plain batched {plain_rate:.2f} substeps/s ({plain_ms:.2f} ms); warm start + batched "
        f"{warm_rate:.2f} substeps/s ({warm_ms:.2f} ms): x{warm_rate / unb_rate:.2f} vs plain unbatched, "
        f"x{warm_rate / plain_rate:.2f} vs plain batched (graph replays {plain_replays}, {warm_replays})")
    jac = None if a.no_jacobi else jacobi_baseline(a, capi, scenes, ctx, ids, r_plain, warm_rate,
                                                   lambda: solver(PLAIN_SWEEPS, False))
    prof = profiled(True, a.steps, warm=warm_k)
    plain_prof = profiled(False, a.steps)
    achieved = prof.tile_alg_bytes / (prof.tile_ms * 1e-3) / 1e9
    plain_achieved = plain_prof.tile_alg_bytes / (plain_prof.tile_ms * 1e-3) / 1e9
    warm_vertex_sweeps = prof.vertex_sweeps / max(1, a.steps)
    copy_ms = ctx.bench_copy(1 << 30, 10)
    copy_gbs = 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    infos = [ctx.info(i) for i in ids]
    log(f"[bench] k_tiles {prof.tile_launches} launches, {prof.tile_ms / max(1, prof.tile_launches):.3f} ms avg, "
        f"{achieved:.0f} GB/s algorithmic (plain run {plain_achieved:.0f}); copy kernel {copy_gbs:.0f} GB/s")
    # HBM bytes per launch from the committed counter runs of the same launch
    # configuration (none if the counters were taken on another one)
    traffic = None
    tf = ROOT / "profiles" / "k_tiles_batch512_traffic.json"
    if tf.exists():
        try:
            key = traffic_key(tuned_warm)
            traffic = json.loads(tf.read_text())["warm"][key]["hbm_bytes_per_launch"]
        except Exception:
            traffic = None
    cpu = None
    if not a.no_cpu:
        vs_rate, desc = cpu_baseline(specs, PLAIN_SWEEPS, a.cpu_seconds)
        cpu = {"value": vs_rate / (nv_total * PLAIN_SWEEPS), "unit": "substeps/s", "cores": 1, "kind": "port",
               "sample": desc, "vertex_sweeps_per_s": vs_rate}
        log(f"[bench] cpu baseline {cpu['value']:.5f} substeps/s ({vs_rate / 1e6:.1f} M vertex-sweeps/s)")
    out = {
        "metric": "cloth substeps/s at fixed residual: warm-start+batched vs plain, 1 MI355X",
        "value": warm_rate,
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": warm_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{len(specs)} grid cloths in one batch (side U{{16..256}}, aspect U(0.5,2), 0-4 random "
                        f"boundary pins, own sphere r U(0.2,0.4)): {nv_total} vertices, {ne_total} distance prims, "
                        f"dt 1/600 s; colour-ordered Gauss-Seidel in LDS tiles",
            "fixed_residual": {"definition": "mean over cloths and substeps of max |d-L|/L at the end of the "
                                             "substep, plain solver with 20 sweeps, same start state",
                               "R": r_plain, "warm_start_residual": r_warm,
                               "substeps_compared": a.match_substeps},
            "warm_start_rule": head,
            "warm_start_other_rule": other,
            "warm_start_without_gate": ungated,
            "plain_substeps_per_s": unb_rate,
            "plain_ms_per_step": unb_ms,
            "plain_definition": "plain solver, unbatched: each cloth stepped alone (its own launches, LDS "
                                "tiles), 20 sweeps, no warm start",
            "speedup_vs_plain": warm_rate / unb_rate,
            "plain_batched_substeps_per_s": plain_rate,
            "plain_batched_ms_per_step": plain_ms,
            "speedup_vs_plain_batched": warm_rate / plain_rate,
            "plain_sweeps": PLAIN_SWEEPS,
            "warm_start_budget_kept": warm_k,
            "warm_start_budget_discarded": PLAIN_SWEEPS,
            "warm_start_kept_fraction": used_frac,
            "warm_start_mean_sweeps": warm_vertex_sweeps / nv_total,
            "warm_start_training": batch_train,
            "warm_start_32x32": scene_info,
            "vertex_iterations_per_s": warm_vertex_sweeps * warm_rate,
            "plain_batched_vertex_iterations_per_s": nv_total * PLAIN_SWEEPS * plain_rate,
            "graph_replays_timed": warm_replays,
            "autotuned": tuned,
            "autotuned_warm": tuned_warm,
            "single_tile_cloths": sum(1 for i in infos if i.whole),
            "tiles": sum(i.ntiles for i in infos),
            "copy_kernel_gbs": copy_gbs,
            "jacobi_baseline": jac,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_tiles",
            "alg_bytes_per_launch": prof.tile_alg_bytes / max(1, prof.tile_launches),
            "avg_launch_ms": prof.tile_ms / max(1, prof.tile_launches),
            "plain_run_achieved": plain_achieved,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
