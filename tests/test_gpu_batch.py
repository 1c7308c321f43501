"""Batch execution modes against the eager batched solver, bit for bit: one
cloth at a time (the unbatched solver), substeps replayed from a captured
graph (each replay checked, with eager work and host reads between
replays), and the autotuner (which must leave results unchanged)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from positionbaseddynamicsold_amd import capi

# two single-tile cloths and one that splits into dependency-cone tiles at
# 2048 vertices per tile; one has a sphere collider
SIZES = [(20, 0), (33, 1), (96, 2)]


def mixed_batch(ctx):
    ids = []
    for n, seed in SIZES:
        cid = ctx.add_grid(n, n, 1.0, 1.0, jitter=1e-4, seed=seed)
        ctx.set_pins(cid, [0, n - 1])
        ids.append(cid)
    ctx.set_params(ids[1], sphere_on=1, sphere_c=(0.5, -0.05, 0.5), sphere_r=0.3)
    ctx.init()
    return ids


def solver(ctx, **kw):
    args = dict(dt=1.0 / 600.0, iterations=7, tile_sweeps=2, check_every=2, tile_max_verts=2048)
    args.update(kw)
    ctx.set_solver(**args)


def assert_same(a, b, ids):
    for i in ids:
        assert np.array_equal(a.get(i, capi.STATE_X), b.get(i, capi.STATE_X)), i
        assert np.array_equal(a.get(i, capi.STATE_V), b.get(i, capi.STATE_V)), i
        assert np.array_equal(a.get(i, capi.STATE_P), a.get(i, capi.STATE_X)), i


@pytest.mark.parametrize("path", [capi.PATH_AUTO, capi.PATH_GLOBAL, capi.PATH_JACOBI])
def test_step_cloth_matches_batch(path):
    print(f"step_cloth path={path}", flush=True)
    with capi.Context() as a, capi.Context() as b:
        ids = mixed_batch(a)
        mixed_batch(b)
        solver(a, path=path)
        solver(b, path=path)
        if path == capi.PATH_AUTO:
            assert [a.info(i).whole for i in ids] == [1, 1, 0]
        for _ in range(5):
            a.step(1)
            for i in ids:
                b.step_cloth(i, 1)
        assert_same(a, b, ids)


def test_step_cloth_with_tolerance_matches_batch():
    # per-cloth stopping: each cloth stops at its own check, batched or not
    with capi.Context() as a, capi.Context() as b:
        ids = mixed_batch(a)
        mixed_batch(b)
        for c in (a, b):
            solver(c, iterations=40, tol=3e-3, check_every=2, tile_sweeps=2)
            c.set_tolerance(ids[0], 1e-3)
        for _ in range(6):
            a.step(1)
            for i in ids:
                b.step_cloth(i, 1)
            assert [a.last_sweeps(i) for i in ids] == [b.last_sweeps(i) for i in ids]
        assert_same(a, b, ids)
        assert len({a.last_sweeps(i) for i in ids}) > 1


def test_step_cloth_rejects_batch_only_modes():
    with capi.Context() as c:
        ids = mixed_batch(c)
        solver(c)
        c.ws_init(0)
        c.ws_enable(True)
        with pytest.raises(capi.PbdError) as e:
            c.step_cloth(ids[0], 1)
        assert e.value.code == capi.PBD_ERR_STATE


@pytest.mark.parametrize("path,ws", [(capi.PATH_AUTO, False), (capi.PATH_AUTO, True), (capi.PATH_GLOBAL, False),
                                     (capi.PATH_JACOBI, False)])
def test_graph_replay_matches_eager(path, ws):
    # 7 sweeps in chunks of 2 end in the other p buffer, so both captured
    # graphs (one per starting buffer) are exercised
    print(f"graph path={path} ws={ws}", flush=True)
    with capi.Context() as g, capi.Context() as e:
        ids = mixed_batch(g)
        mixed_batch(e)
        solver(g, path=path, graph=1)
        solver(e, path=path, graph=0)
        if ws:
            for c in (g, e):
                c.ws_init(3)
                c.ws_enable(True)
        for rep in range(6):
            g.step(1)
            e.step(1)
            g.sync()
            assert_same(g, e, ids)
            if ws:
                assert [g.ws_last_used(i) for i in ids] == [e.ws_last_used(i) for i in ids]
            # eager work between replays: a residual reduction and a state read
            g.residual_all()
        g.step(4)
        e.step(4)
        assert_same(g, e, ids)
        assert g.profile_read().graph_replays == 10
        assert e.profile_read().graph_replays == 0
        # a state write between replays is seen by the next replay
        x = e.get(ids[2])
        x[:, 1] += np.float32(0.01)
        g.set(ids[2], capi.STATE_X, x)
        e.set(ids[2], capi.STATE_X, x)
        g.step(2)
        e.step(2)
        assert_same(g, e, ids)


def test_graph_not_used_with_tolerance_or_profiling():
    with capi.Context() as g:
        mixed_batch(g)
        solver(g, graph=1, tol=1e-3)
        g.step(2)
        assert g.profile_read().graph_replays == 0
        solver(g, graph=1)
        g.profile(True)
        g.step(2)
        assert g.profile_read().graph_replays == 0
        g.profile(False)
        g.step(2)
        assert g.profile_read().graph_replays == 2


def test_autotune_leaves_results_unchanged():
    with capi.Context() as a, capi.Context() as b:
        ids = mixed_batch(a)
        mixed_batch(b)
        solver(a, tile_max_verts=0, graph=1)
        solver(b, tile_max_verts=0, graph=1)
        a.step(2)
        b.step(2)
        p, bits, ms, n = a.autotune(1)
        print(f"autotune: threads {p.block_threads} tile_sweeps {p.tile_sweeps} max_verts {p.tile_max_verts} "
              f"{ms:.3f} ms/substep over {n} trials", flush=True)
        # threads x tile_sweeps {1, 2, 4, 5} x tile sizes (the whole LDS, half of
        # it, 4096 and 2048 vertices), then the other way of stepping
        # single-tile cloths (split_whole) at the best of those
        assert n == 3 * 4 * 4 + 1
        assert bits == -1 and ms > 0.0
        assert p.block_threads in (256, 512, 1024) and p.tile_sweeps in (1, 2, 4, 5)
        assert p.iterations == 7 and p.graph == 1
        # the trials leave the state where it was
        assert_same(a, b, ids)
        a.step(5)
        b.step(5)
        assert_same(a, b, ids)


@pytest.mark.parametrize("mode", ["plain", "warm", "tol"])
def test_split_whole_matches(mode):
    # single-tile cloths stepped a few sweeps per launch (split_whole = 1) give
    # the same positions, sweep counts and fallback decisions as one launch each
    print(f"split_whole mode={mode}", flush=True)
    with capi.Context() as a, capi.Context() as b:
        ids = mixed_batch(a)
        mixed_batch(b)
        kw = dict(iterations=40, tol=3e-3, check_every=2) if mode == "tol" else {}
        solver(a, graph=1, **kw)
        solver(b, graph=1, split_whole=1, **kw)
        assert [a.info(i).whole for i in ids] == [1, 1, 0]
        assert [b.info(i).whole for i in ids] == [0, 0, 0]
        if mode == "warm":
            for c in (a, b):
                c.ws_init(3)
                c.ws_enable(True)
                c.set_solver(dt=1.0 / 600.0, iterations=7, tile_sweeps=2, check_every=2, tile_max_verts=2048,
                             graph=1, warm_iterations=3, split_whole=1 if c is b else 0)
        for _ in range(6):
            a.step(1)
            b.step(1)
            assert_same(a, b, ids)
            if mode == "warm":
                assert [a.ws_last_used(i) for i in ids] == [b.ws_last_used(i) for i in ids]
            if mode == "tol":
                assert [a.last_sweeps(i) for i in ids] == [b.last_sweeps(i) for i in ids]
