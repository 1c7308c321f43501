"""The 512-cloth batch (side U{16..256}, aspect U(0.5, 2), 0-4 random pins,
own sphere each; 15.4 M vertices) at its own scale, warm start and fallback
on, substeps replayed from captured graphs.

Reference semantics: one cloth per SolverPBD (PBD_Basic.cuh:224-233), stepped
by the driver loop of PositionBasedDynamics.cpp:263-290.  So each cloth of the
batch must equal that cloth stepped alone by the oracle, whatever else the
batch holds, and a cloth whose prediction the fallback discards must equal
the plain batched solver's substep.

Checked, per substep, for two sweeps-per-launch settings (the cone tiles of
the multi-tile cloths differ between them):
  - every discarded cloth is bit-identical to the plain batched solver
    (20 sweeps) started from the same state;
  - a sample of cloths covering the tile regimes (single-tile, few tiles,
    many tiles; kept and discarded; pinned and free) equals the oracle stepped
    alone: advect, plus the GPU's predicted correction when kept, then the
    kept (3) or plain (20) sweeps with the sphere after each sweep, integrate;
  - a second run from the same state is bit-identical for every cloth, and
    so is the eager (uncaptured) substep.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from positionbaseddynamicsold_amd import capi, scenes

PLAIN, WARM = 20, 3


@pytest.fixture(scope="module")
def batch():
    print("config5: building 512 cloths", flush=True)
    specs = scenes.config5(seed=0, count=512)
    ctx = capi.Context(0)
    ids = [scenes.add_to_context(ctx, s) for s in specs]
    ctx.init()
    ctx.set_solver(dt=scenes.DT, iterations=PLAIN, tile_sweeps=4, check_every=4, graph=1)
    ctx.step(3)
    # a short training run on the batch's own samples: the test is about the
    # substep's bit-exactness under a mix of kept and discarded predictions
    ctx.ws_init(0)
    ctx.ws_collect(40, 1)
    loss = ctx.ws_train(epochs=1, batch=4096, lr=1e-3, seed=1)
    print(f"config5: {ctx.ws_num_samples()} samples, loss {loss:.3f}", flush=True)
    ctx.step(1)
    ctx.sync()
    ctx.snapshot()
    info = [ctx.info(i) for i in ids]
    off = np.concatenate([[0], np.cumsum([i.nv for i in info])])
    yield ctx, ids, specs, info, off
    ctx.close()


def set_solver(ctx, ts, ws, graph=1):
    ctx.ws_enable(ws)
    ctx.set_solver(dt=scenes.DT, iterations=PLAIN, tile_sweeps=ts, check_every=ts, graph=graph,
                   warm_iterations=WARM if ws else 0, warm_rule=1, warm_gate=1e-4)


def state(ctx, off):
    """(x, v) of every cloth, from one whole-batch read."""
    x, v = ctx.state_save()
    return [(x[a:b], v[a:b]) for a, b in zip(off[:-1], off[1:])]


def pick_sample(info, specs, used):
    """The smallest cloth of each (regime, kept, pinned) class that the batch
    has -- regimes by tile count: single-tile, few tiles (2-8), many tiles --
    plus the cloth with the most tiles."""
    def regime(k):
        return "whole" if info[k].whole else ("few" if info[k].ntiles <= 8 else "many")
    chosen, seen = [], set()
    for k in sorted(range(len(info)), key=lambda k: info[k].nv):
        key = (regime(k), used[k], bool(specs[k].pins))
        if key not in seen:
            seen.add(key)
            chosen.append(k)
    biggest = max(range(len(info)), key=lambda k: info[k].ntiles)
    if biggest not in chosen:
        chosen.append(biggest)
    return chosen


_ORACLE = {}


def oracle_substep(orc, k, spec, x, v, delta, sweeps):
    if k not in _ORACLE:
        pos, tris = orc.grid_mesh(spec.rx, spec.ry, spec.sx, spec.sy, jitter=spec.jitter, seed=spec.seed)
        _ORACLE[k] = orc.OracleCloth(pos, tris, orc.grid_pairs(spec.rx, spec.ry), pins=spec.pins, iterations=PLAIN)
    oc = _ORACLE[k]
    prm = oc.params(dt=scenes.DT, sphere=spec.sphere)
    oc.x[:] = x
    oc.v[:] = v
    oc.advect(prm)
    if delta is not None:
        oc.p[:] = oc.p + delta  # float32 adds, as the tile kernel's first load
    r = C.c_double()
    n = oc.L.orc_substep(C.byref(oc.c), C.byref(prm), sweeps, 1, -1.0, C.byref(r))
    assert n == sweeps
    return oc.x.copy(), oc.v.copy()


@pytest.mark.parametrize("ts", [4, 2])
def test_config5_warm_graph_against_plain_and_oracle(batch, orc, ts):
    ctx, ids, specs, info, off = batch
    nwhole = sum(i.whole for i in info)
    print(f"config5 ts={ts}: {len(ids)} cloths, {sum(i.nv for i in info)} vertices, {nwhole} single-tile, "
          f"{sum(i.ntiles for i in info)} tiles", flush=True)
    assert 0 < nwhole < len(ids)
    ctx.snapshot(restore=True)
    sample = None
    for sub in range(2):
        x0, v0 = ctx.state_save()
        start = [(x0[a:b], v0[a:b]) for a, b in zip(off[:-1], off[1:])]
        # plain batched substep from this state
        set_solver(ctx, ts, False)
        ctx.step(1)
        plain = state(ctx, off)
        # warm-started substep from the same state (graph replay)
        ctx.state_load(x0, v0)
        set_solver(ctx, ts, True)
        ctx.step(1)
        warm = state(ctx, off)
        used = [ctx.ws_last_used(i) for i in ids]
        sweeps = [ctx.last_sweeps(i) for i in ids]
        kept = sum(used)
        print(f"config5 ts={ts} substep {sub}: kept {kept} / {len(ids)}", flush=True)
        assert 0 < kept < len(ids)
        for k, i in enumerate(ids):
            assert sweeps[k] == (WARM if used[k] else PLAIN), k
            if not used[k]:
                assert np.array_equal(warm[k][0], plain[k][0]), f"cloth {k} x"
                assert np.array_equal(warm[k][1], plain[k][1]), f"cloth {k} v"
        if sample is None:
            sample = pick_sample(info, specs, used)
            print(f"config5 sample (cloth, nv, tiles, kept, pins): "
                  f"{[(k, info[k].nv, info[k].ntiles, used[k], len(specs[k].pins)) for k in sample]}", flush=True)
            assert len(sample) >= 6
            assert {info[k].whole for k in sample} == {0, 1}
        for k in sample:
            delta = ctx.ws_last_delta(ids[k]) if used[k] else None
            ox, ov = oracle_substep(orc, k, specs[k], start[k][0], start[k][1], delta, WARM if used[k] else PLAIN)
            assert np.array_equal(warm[k][0], ox), f"cloth {k} x against the oracle"
            assert np.array_equal(warm[k][1], ov), f"cloth {k} v against the oracle"
        # the same warm substep again: replayed and eager, bit-identical
        ctx.state_load(x0, v0)
        ctx.step(1)
        again = state(ctx, off)
        ctx.state_load(x0, v0)
        set_solver(ctx, ts, True, graph=0)
        ctx.step(1)
        eager = state(ctx, off)
        assert [ctx.ws_last_used(i) for i in ids] == used
        for k in range(len(ids)):
            for a, b, c in ((warm[k][0], again[k][0], eager[k][0]), (warm[k][1], again[k][1], eager[k][1])):
                assert np.array_equal(a, b), f"cloth {k}: replay differs"
                assert np.array_equal(a, c), f"cloth {k}: eager differs"
