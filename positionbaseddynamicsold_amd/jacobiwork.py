"""The Jacobi path on the 512-cloth batch as a short program for kernel traces.

    python -m positionbaseddynamicsold_amd.jacobiwork [--sweeps S] [--substeps K] [--omega W]

Builds the batch bench.py measures (512 grid cloths, 15.4 M vertices), runs
two untimed Jacobi substeps and then K timed ones of S sweeps each, replayed
from a captured graph, and prints one JSON line: sizes, the algorithmic bytes
of one k_jacobi sweep (48 B per vertex, 8 B per distance incidence, 24 B per
bending incidence) and the wall time per substep.
"""
from __future__ import annotations

import argparse
import json
import time

from . import capi, scenes


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sweeps", type=int, default=10)
    ap.add_argument("--substeps", type=int, default=10)
    ap.add_argument("--omega", type=float, default=1.5)
    ap.add_argument("--count", type=int, default=512)
    a = ap.parse_args()
    specs = scenes.config5(seed=0, count=a.count)
    ctx = capi.Context(0)
    ids = [scenes.add_to_context(ctx, s) for s in specs]
    ctx.init()
    ctx.set_solver(dt=scenes.DT, iterations=a.sweeps, path=capi.PATH_JACOBI, jacobi_omega=a.omega, graph=1)
    ctx.step(2)
    ctx.sync()
    t = time.perf_counter()
    ctx.step(a.substeps)
    ctx.sync()
    ms = (time.perf_counter() - t) * 1e3 / a.substeps
    infos = [ctx.info(i) for i in ids]
    nv = sum(i.nv for i in infos)
    out = {"cloths": len(ids), "vertices": nv, "distance_prims": sum(i.ne for i in infos),
           "sweeps": a.sweeps, "substeps": a.substeps, "omega": a.omega, "ms_per_substep": ms,
           "alg_bytes_per_sweep": 48 * nv + 8 * sum(2 * i.ne for i in infos) + 24 * sum(4 * i.nb for i in infos)}
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
