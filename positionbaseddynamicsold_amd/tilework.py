"""The bench's tile-kernel passes on the 512-cloth batch, without the training
around them, as a short program for hardware-counter collection.

    python -m positionbaseddynamicsold_amd.tilework --network NET.npz [--steps K] [--warm-tile-sweeps N]

NET.npz is the batch network bench.py writes with --save-network (parameters,
normalisation, warm budget, fallback rule, tile configuration).  The program
rebuilds the batch (same seeded generator), repeats the bench's substeps up to
its snapshot (warm-up, sample-recording substeps with their sweep count,
settling substeps), then runs the two eager passes bench.py times with HIP
events -- one untimed and K timed warm-started substeps, then the same for
the plain solver -- and the 1 GiB calibration copy, and prints one JSON line
with the tile kernel's average launch time per pass.  It launches a few
thousand kernels, against the bench's hundreds of thousands, so that
rocprofv3 --pmc runs stay short; tracestats --fetch/--write turns the counters
of those passes into HBM bytes per launch.
"""
from __future__ import annotations

import argparse
import json

import numpy as np

from . import capi, scenes


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--network", required=True)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warm-tile-sweeps", type=int, default=0,
                    help="sweeps per launch of the warm-started pass (default: the network file's)")
    a = ap.parse_args()
    net = np.load(a.network)
    cfg = {k: int(net[k]) for k in ("count", "seed", "warmup", "collect_substeps", "target_sweeps", "settle",
                                     "plain_sweeps", "warm_k", "warm_rule", "tile_sweeps", "tile_max_verts",
                                     "block_threads")}
    cfg["split_whole"] = int(net["split_whole"]) if "split_whole" in net else 0
    # the warm-started run's own launch configuration, when the network file holds one
    wcfg = {k: int(net["warm_" + k]) if "warm_" + k in net else cfg[k]
            for k in ("tile_sweeps", "tile_max_verts", "block_threads", "split_whole")}
    if a.warm_tile_sweeps > 0:
        wcfg["tile_sweeps"] = a.warm_tile_sweeps
    specs = scenes.config5(seed=cfg["seed"], count=cfg["count"])
    ctx = capi.Context(0)
    for s in specs:
        scenes.add_to_context(ctx, s)
    ctx.init()

    def solver(ws, graph=0, warm=0):
        c = wcfg if ws else cfg
        ctx.ws_enable(ws)
        ctx.set_solver(dt=scenes.DT, iterations=cfg["plain_sweeps"], check_every=c["tile_sweeps"],
                       tile_sweeps=c["tile_sweeps"], tile_max_verts=c["tile_max_verts"],
                       block_threads=c["block_threads"], graph=graph, warm_iterations=warm,
                       warm_rule=cfg["warm_rule"], warm_gate=float(net["warm_gate"]) if "warm_gate" in net else 0.0,
                       split_whole=c["split_whole"])

    # the bench's substeps before its snapshot
    solver(False, graph=1)
    ctx.step(2)
    ctx.step(cfg["warmup"])
    ctx.ws_init(0)
    for _ in range(cfg["collect_substeps"]):
        ctx.ws_collect(cfg["target_sweeps"], 1)
        ctx.step(1)
    ctx.ws_set_params(net["params"])
    ctx.ws_set_norm(net["norm"])
    solver(False, graph=1)
    ctx.step(cfg["settle"])
    ctx.sync()
    ctx.snapshot()

    def profiled(ws, warm=0):
        ctx.snapshot(restore=True)
        solver(ws, warm=warm)
        ctx.step(1)
        ctx.sync()
        ctx.profile(True)
        ctx.profile_read(reset=True)
        ctx.step(a.steps)
        ctx.sync()
        prof = ctx.profile_read(reset=True)
        ctx.profile(False)
        return {"launches": prof.tile_launches, "avg_launch_ms": prof.tile_ms / max(1, prof.tile_launches),
                "alg_bytes_per_launch": prof.tile_alg_bytes / max(1, prof.tile_launches),
                "achieved_gbs": prof.tile_alg_bytes / (prof.tile_ms * 1e-3) / 1e9}

    out = {"warm": profiled(True, warm=cfg["warm_k"]), "plain": profiled(False),
           "warm_config": wcfg,
           "plain_config": {k: cfg[k] for k in ("tile_sweeps", "tile_max_verts", "block_threads", "split_whole")}}
    out["copy_ms"] = ctx.bench_copy(1 << 30, 10)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
