"""The warm start's one-launch tile work against the same substeps run as
one launch per chunk (per_chunk_launches = 1).  Two queued forms: 0, where
a separate residual kernel decides and a device-built queue follows, and 2,
where the launch itself computes the fallback residuals and keeps or
discards each cloth's prediction (fixed first-chunk items, later chunks
appended by the tile that completes its cloth's decision).  In both, workgroups take
(tile, chunk) items and a tile's later chunk waits only for its neighbour
tiles' previous chunk.  Checked: bit-identical states and keep decisions on
an uneven batch with kept and discarded multi-tile cloths, with and without
bending prims, at several chunk lengths and warm budgets (including a budget
longer than one chunk, so that kept cloths also have later chunks), graph
replay and eager; the deciding form ran; no bounded wait gave up."""
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
@pytest.mark.parametrize("form", [0, 2])
def test_queued_launch_equals_per_chunk_launches(batch, form, ts, warm, max_verts, graph):
    max_verts *= batch.budget_scale
    print(f"queue form {form} vs per-chunk: ts {ts} warm {warm} max_verts {max_verts} graph {graph}", flush=True)
    x0, v0, k0 = run(batch, 1, ts, warm, graph, max_verts)
    before = batch.profile_read(reset=False).queue_deciding_launches  # counted on the device
    x1, v1, k1 = run(batch, form, ts, warm, graph, max_verts)
    prof = batch.profile_read(reset=False)
    deciding = prof.queue_deciding_launches - before
    print(f"  kept per substep {k1}, deciding launches {deciding}, queue wait timeouts {prof.queue_wait_timeouts}",
          flush=True)
    assert deciding == (6 if form == 2 else 0)
    assert k0 == k1
    assert 0 < min(k1) and max(k1) < 40  # both kept and discarded cloths
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1)
    assert prof.queue_wait_timeouts == 0
