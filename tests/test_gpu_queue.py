"""The warm start's one-launch tile work (per_chunk_launches = 0: workgroups
walk a device-built queue of (tile, chunk) items, a tile's later chunk
waiting only for its neighbour tiles' previous chunk) against the same
substeps run as one launch per chunk: bit-identical states on an uneven
batch with kept and discarded multi-tile cloths, at several chunk lengths
and warm budgets (including a budget longer than one chunk, so that kept
cloths also have later chunks), graph replay and eager, and no wait of the
queued launch ending on its bounded spin instead of the neighbour's
chunk."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from positionbaseddynamicsold_amd import capi, scenes


@pytest.fixture(scope="module", params=[False, True], ids=["distance", "bending"])
def batch(request):
    # the bending batch (dihedral prims at stiffness 0.1) runs the bending
    # form of the tile kernels
    specs = scenes.config5(seed=7, count=40)
    ctx = capi.Context()
    for s in specs:
        scenes.add_to_context(ctx, s, bending=request.param, stiffness=(1.0, 0.1, 1.0))
    ctx.init()
    ctx.set_solver(dt=scenes.DT, iterations=20)
    ctx.step(12)
    ctx.ws_init(0)
    ctx.ws_collect(60, 1)
    ctx.ws_train(epochs=8, batch=4096, lr=1e-3, seed=1)
    ctx.sync()
    ctx.snapshot()
    # a dihedral prim spans two triangles, so a vertex's dependency cone over
    # the same sweeps is several times larger: the bending batch gets four
    # times the tile budget (0 keeps the default budget)
    ctx.budget_scale = 4 if request.param else 1
    yield ctx
    ctx.close()


def run(ctx, per_chunk, ts, warm, graph, max_verts, steps=6):
    ctx.snapshot(restore=True)
    ctx.ws_enable(True)
    ctx.set_solver(dt=scenes.DT, iterations=20, check_every=ts, tile_sweeps=ts, tile_max_verts=max_verts,
                   graph=graph, warm_iterations=warm, warm_rule=1, warm_gate=1e-4, per_chunk_launches=per_chunk)
    kept = []
    for _ in range(steps):
        ctx.step(1)
        kept.append(sum(ctx.ws_last_used(i) for i in range(40)))
    x, v = ctx.state_save()
    return x.copy(), v.copy(), kept


@pytest.mark.parametrize("ts,warm,max_verts", [(2, 1, 2048), (1, 1, 1024), (2, 3, 2048), (3, 1, 0)])
@pytest.mark.parametrize("graph", [1, 0])
def test_queued_launch_equals_per_chunk_launches(batch, ts, warm, max_verts, graph):
    max_verts *= batch.budget_scale
    print(f"queue vs per-chunk: ts {ts} warm {warm} max_verts {max_verts} graph {graph}", flush=True)
    x0, v0, k0 = run(batch, 1, ts, warm, graph, max_verts)
    x1, v1, k1 = run(batch, 0, ts, warm, graph, max_verts)
    prof = batch.profile_read(reset=False)  # the timeout count is read from the device either way
    print(f"  kept per substep {k1}, queue wait timeouts {prof.queue_wait_timeouts}", flush=True)
    assert k0 == k1
    assert 0 < min(k1) and max(k1) < 40  # both kept and discarded cloths
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1)
    assert prof.queue_wait_timeouts == 0
