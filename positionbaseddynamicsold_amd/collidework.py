"""The collision path of config 3 (or 2) at full size as a short program for
kernel traces and hardware counters.

    python -m positionbaseddynamicsold_amd.collidework [--config 3] [--mode robust|reference]
        [--edge-edge] [--settle N] [--substeps K]

Builds the scene (config 3: the 512 x 512 folded cloth with self contact over
the 0.4 m box mesh; config 2: the 256 x 256 cloth with bending over the
icosphere), runs N substeps until contact, then K substeps with the contact
counters reset, copies 1 GiB with the calibration copy kernel, and prints one
JSON line: scene sizes, hash table, and per-run totals of detections,
triangles hashed, cell entries, raw and distinct candidate pairs, contacts,
response passes (and the edge-edge candidates and contacts with --edge-edge),
plus the wall time per substep.  positionbaseddynamicsold_amd
.tracestats --collide turns a kernel trace (and FETCH_SIZE / WRITE_SIZE
counter runs) of it into per-kernel time and HBM rate.
"""
from __future__ import annotations

import argparse
import json
import time
from dataclasses import replace

from . import capi, scenes


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=int, default=3, choices=(2, 3))
    ap.add_argument("--mode", default="robust", choices=("robust", "reference"))
    ap.add_argument("--settle", type=int, default=0, help="substeps before the counted ones (a kernel trace "
                    "then holds more detections than the counts)")
    ap.add_argument("--substeps", type=int, default=24)
    ap.add_argument("--edge-edge", action="store_true", help="edge-edge contacts too")
    a = ap.parse_args()
    sc = scenes.config3() if a.config == 3 else scenes.config2()
    if a.mode == "robust":
        sc = sc.robust()
    if a.edge_edge:
        sc = replace(sc, edge_edge=1)
    ctx = capi.Context(0)
    cid, sid, cell = scenes.add_ccd_scene(ctx, sc)
    if a.settle > 0:
        ctx.step(a.settle)
    ctx.sync()
    ctx.profile(True)
    ctx.profile_read(reset=True)
    t = time.perf_counter()
    ctx.step(a.substeps)
    ctx.sync()
    ms = (time.perf_counter() - t) * 1e3 / max(1, a.substeps)
    prof = ctx.profile_read(reset=True)
    ctx.profile(False)
    copy_ms = ctx.bench_copy(1 << 30, 10)
    ci, si = ctx.info(cid), ctx.info(sid)
    out = {
        "config": a.config, "mode": a.mode, "settle": a.settle, "substeps": a.substeps,
        "cloth_vertices": ci.nv, "cloth_triangles": ci.nt, "static_vertices": si.nv, "static_triangles": si.nt,
        "distance_prims": ci.ne, "bending_prims": ci.nb, "sweeps": sc.iterations, "detect_every": sc.detect_every,
        "resolve_passes": sc.resolve_passes, "table_bits": sc.table_bits, "cell": cell,
        "detections": prof.ccd_detections, "triangles": prof.ccd_triangles, "cell_entries": prof.ccd_cell_entries,
        "raw_pairs": prof.ccd_raw_pairs, "pairs": prof.ccd_pairs, "contacts": prof.ccd_contacts,
        "resolved": prof.ccd_resolved, "edge_edge": int(a.edge_edge), "ee_raw_pairs": prof.ccd_ee_raw_pairs,
        "ee_pairs": prof.ccd_ee_pairs, "ee_contacts": prof.ccd_ee_contacts,
        "ms_per_substep": ms, "copy_ms": copy_ms,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
