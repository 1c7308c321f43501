"""Summaries of rocprofv3 CSV output of bench.py.

    python -m positionbaseddynamicsold_amd.tracestats [TRACE.csv] [--steps K]
        [--fetch FETCH_counter_collection.csv --write WRITE_counter_collection.csv --passes PASSES.json]

TRACE.csv is `rocprofv3 --kernel-trace --output-format csv` output
(`*_kernel_trace.csv`).  Prints per-kernel statistics for the whole run, then
for the two eager passes bench.py times with HIP events after its headline
runs -- K warm-started substeps followed by K plain substeps, each preceded
by one untimed substep, just before the copy-kernel calibration -- so the
tile kernel's average launch duration can be compared with the bench's own
figure.  Substeps are delimited by the per-substep k_final_buffers launch (the
tile path's last kernel of a projection; integrate may run inside the tiles).

With counter files from `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE`
runs (kernel filter "k_tiles|k_copy|k_final_buffers") of `bench.py
--passes-only`, which repeats the same two passes, it prints the tile
kernel's HBM traffic per launch in those passes, scaled by what the counters
read for the calibration copy (1 GiB read + 1 GiB written per launch), next
to the bytes the launches must move as designed (the JSON line that run
printed, --passes).  With --collide (the JSON line of
positionbaseddynamicsold_amd.collidework), it summarises that program's
trace instead: every kernel's time and, with counter files of the same
program, HBM bytes and GB/s per kernel, plus the algorithmic rate of the
kernels whose bytes per unit are fixed (hash keys 56 B per triangle,
narrowphase 104 B per candidate, edge-edge narrowphase 152 B per candidate,
response 160 B per contact and pass).
"""
from __future__ import annotations

import argparse
import csv
import json
import statistics
from collections import defaultdict

COPY_BYTES = 1 << 30  # bench.py calibrates with pbd_bench_copy(1 GiB)


def short(name: str) -> str:
    return name.split("(")[0].replace("pbd::", "")


def load_trace(path):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    return [(short(r["Kernel_Name"]), int(r["Start_Timestamp"]), int(r["End_Timestamp"]), int(r["Grid_Size_X"]))
            for r in rows]


def passes(seq, steps, name_of=lambda e: e[0], grid_of=lambda e: e[3]):
    """Index ranges [a, b) of the warm and plain profiled passes in seq."""
    copies = [i for i, e in enumerate(seq) if name_of(e) == "k_copy"]
    big = max(grid_of(seq[i]) for i in copies)
    c0 = max(i for i in copies if grid_of(seq[i]) == big)
    while c0 > 0 and name_of(seq[c0 - 1]) == "k_copy" and grid_of(seq[c0 - 1]) == big:
        c0 -= 1
    # the last kernel of each substep: k_final_buffers (plain, per-chunk
    # launches) or the queued tile launch (warm start, which records the
    # final buffers in its schedule kernel)
    ints = [i for i in range(c0) if name_of(seq[i]) in ("k_final_buffers", "k_tiles_queue")]
    need = 2 * (steps + 1)
    if len(ints) < need:
        raise SystemExit("trace does not hold the two profiled passes")
    ends = ints[-need:]
    # pass = substeps 1..steps of each group of steps + 1 (the first is untimed)
    warm = (ends[0] + 1, ends[steps] + 1)
    plain = (ends[steps + 1] + 1, ends[2 * steps + 1] + 1)
    return warm, plain, c0


def table(events, title):
    per = defaultdict(list)
    for n, s, e, _ in events:
        per[n].append((e - s) / 1e3)
    tot = sum(sum(v) for v in per.values())
    print(f"\n{title}")
    print(f"{'kernel':28s} {'calls':>7s} {'total ms':>10s} {'avg us':>10s} {'share':>6s}")
    for n, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        print(f"{n:28s} {len(v):7d} {sum(v) / 1e3:10.3f} {statistics.fmean(v):10.1f} {100 * sum(v) / tot:5.1f}%")
    return per


def counters(path):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    out = []
    for r in rows:
        grid = int(r.get("Grid_Size", r.get("Grid_Size_X", 0)) or 0)
        out.append((short(r["Kernel_Name"]), float(r["Counter_Value"]), grid))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace", nargs="?")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--passes", help="JSON line printed by the counted bench.py --passes-only run")
    ap.add_argument("--collide", help="JSON line printed by the collidework run")
    a = ap.parse_args()
    if a.collide:
        collide_summary(a.trace, a.collide, a.fetch, a.write)
        return
    if a.trace:
        ev = load_trace(a.trace)
        table(ev, "whole run (all launches)")
        (wa, wb), (pa, pb), _ = passes(ev, a.steps)
        w = table(ev[wa:wb], f"profiled warm-started pass ({a.steps} substeps)")
        p = table(ev[pa:pb], f"profiled plain pass ({a.steps} substeps)")
        for label, per in (("warm", w), ("plain", p)):
            # the tile kernel in any of its forms (per-chunk launches, the queued launch, with bending)
            t = [d for n, v in per.items() if n.startswith("k_tiles") for d in v]
            if t:
                print(f"k_tiles {label}: {len(t)} launches, average {statistics.fmean(t) / 1e3:.4f} ms")
    if a.fetch and a.write:
        for label, rng in (("warm", 0), ("plain", 1)):
            vals = []
            for path in (a.fetch, a.write):
                seq = counters(path)
                (wa2, wb2), (pa2, pb2), c0 = passes(seq, a.steps, grid_of=lambda e: e[2])
                cal = statistics.fmean(v for n, v, g in seq[c0:] if n == "k_copy")
                a2, b2 = (wa2, wb2) if rng == 0 else (pa2, pb2)
                tiles = [v for n, v, _ in seq[a2:b2] if n.startswith("k_tiles")]
                vals.append((statistics.fmean(tiles), cal))
            (f, fc), (wr, wc) = vals
            scale_f = COPY_BYTES / (fc * 1024.0)
            scale_w = COPY_BYTES / (wc * 1024.0)
            traffic = f * 1024.0 * scale_f + wr * 1024.0 * scale_w
            print(f"k_tiles {label}: FETCH_SIZE {f:.0f} KiB, WRITE_SIZE {wr:.0f} KiB per launch; calibration copy "
                  f"reads {fc:.0f} KiB / writes {wc:.0f} KiB for 1 GiB each (scale {scale_f:.4f}, {scale_w:.4f}); "
                  f"HBM traffic {traffic / 1e9:.4f} GB per launch")
            if a.passes:
                run = json.loads(open(a.passes).read().strip().splitlines()[-1])[label]
                need = run["bytes_per_launch"]
                print(f"  bytes the launches must move as designed: {need / 1e9:.4f} GB per launch "
                      f"(traffic / design {traffic / need:.3f}); under the counters the launches averaged "
                      f"{run['avg_launch_ms']:.4f} ms: {traffic / (run['avg_launch_ms'] * 1e-3) / 1e9:.0f} GB/s of "
                      f"HBM traffic, {need / (run['avg_launch_ms'] * 1e-3) / 1e9:.0f} GB/s by design bytes")


HBM_PEAK = 8.0e12

# algorithmic bytes per unit (work-file count) of the kernels with a fixed per-unit figure
COLLIDE_ALG = {"k_tri_keys": (56, "triangles"), "k_narrow": (104, "pairs"), "k_resolve_acc": (160, "resolved"),
               "k_narrow_ee": (152, "ee_pairs")}


def collide_summary(trace, work_path, fetch=None, write=None):
    work = json.loads(open(work_path).read().strip().splitlines()[-1])
    ev = load_trace(trace)
    per = defaultdict(list)
    for n, s, e, _ in ev:
        per[n].append((e - s) / 1e3)
    tot = sum(sum(v) for v in per.values())
    hbm = {}
    if fetch and write:
        for path, kind in ((fetch, "f"), (write, "w")):
            seq = counters(path)
            copies = [v for n, v, g in seq if n == "k_copy"]
            cal = statistics.fmean(copies[-10:])  # the 1 GiB calibration copies end the run
            scale = COPY_BYTES / (cal * 1024.0)
            for n, v, _ in seq:
                if n == "k_copy":
                    continue
                hbm.setdefault(n, {"f": 0.0, "w": 0.0})[kind] += v * 1024.0 * scale
    print(f"config {work['config']} ({work['mode']} contact handling): {work['cloth_vertices']} cloth vertices, "
          f"{work['cloth_triangles'] + work['static_triangles']} triangles, table 2^{work['table_bits']}, "
          f"{work['substeps']} substeps x {work['sweeps']} sweeps, {work['detections']} detections")
    if work.get("edge_edge"):
        print(f"edge-edge per detection: {work['ee_raw_pairs'] / max(1, work['detections']):.0f} raw pairs, "
              f"{work['ee_pairs'] / max(1, work['detections']):.0f} distinct pairs, "
              f"{work['ee_contacts'] / max(1, work['detections']):.0f} contacts")
    print(f"per detection: {work['cell_entries'] / max(1, work['detections']):.0f} cell entries, "
          f"{work['raw_pairs'] / max(1, work['detections']):.0f} raw pairs, "
          f"{work['pairs'] / max(1, work['detections']):.0f} distinct pairs, "
          f"{work['contacts'] / max(1, work['detections']):.0f} contacts; {work['ms_per_substep']:.2f} ms per substep")
    print(f"\n{'kernel':26s} {'calls':>6s} {'total ms':>9s} {'share':>6s} {'HBM GB':>8s} {'HBM GB/s':>9s} "
          f"{'of peak':>7s} {'alg GB/s':>9s}")
    for n, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        t = sum(v) * 1e-6  # durations are in microseconds
        line = f"{n:26s} {len(v):6d} {sum(v) / 1e3:9.3f} {100 * sum(v) / tot:5.1f}%"
        if n in hbm:
            b = hbm[n]["f"] + hbm[n]["w"]
            line += f" {b / 1e9:8.3f} {b / t / 1e9:9.0f} {b / t / HBM_PEAK:6.1%}"
        else:
            line += f" {'':8s} {'':9s} {'':7s}"
        if n in COLLIDE_ALG:
            per_unit, key = COLLIDE_ALG[n]
            line += f" {per_unit * work[key] / t / 1e9:9.0f}"
        print(line)


if __name__ == "__main__":
    main()
