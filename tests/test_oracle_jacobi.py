"""Pins of the oracle's Jacobi variant (orc_jacobi_sweep / orc_substep_jacobi):
constraint averaging (Macklin et al. 2014, sec. 4.2) over the prims' own
projections.

- A float64 numpy restatement from the formulas (distance correction
  -w_i/(w_i+w_j) (d-L) (p_i-p_j)/d k, averaged with omega/n over each
  vertex's prims) agrees to float32 rounding on a pinned, stretched grid.
- With bending, each prim's correction comes from the oracle's single-prim
  projection (pinned elsewhere) and numpy does the averaging.
- One prim alone with omega = 1 and k = 1 satisfies its constraint.
- A cloth at rest does not move; pinned vertices never move.
- The result does not depend on the vertex numbering (up to summation order).
- The residual falls over sweeps.
"""
import numpy as np
import pytest

DT = 1.0 / 600.0


def stretched_grid(orc, rx=7, ry=6, seed=0, bending=False, stiffness=(0.8, 0.5, 1.0)):
    pos, tris = orc.grid_mesh(rx, ry, 1.0, 0.8, jitter=1e-3, seed=seed)
    pairs = orc.grid_pairs(rx, ry)
    oc = orc.OracleCloth(pos, tris, pairs, pins=[0, rx - 1], bending=bending, stiffness=stiffness)
    rng = np.random.default_rng(seed + 10)
    p = pos + rng.normal(scale=0.02, size=pos.shape).astype(np.float32)
    p[[0, rx - 1]] = pos[[0, rx - 1]]
    return oc, p.astype(np.float32)


def numpy_jacobi(oc, p, sweep, omega, bend_fn=None):
    """One averaged Jacobi sweep, float64, from the formulas."""
    p = p.astype(np.float64)
    nv = len(p)
    acc = np.zeros_like(p)
    cnt = np.zeros(nv, np.int64)
    kd = float(oc.kd[min(sweep, oc.kmax - 1)])
    w = oc.w.astype(np.float64)
    for prim, (st, n) in enumerate(oc.prim_list):
        ids = oc.indices[st:st + n]
        cnt[ids] += 1
        if n == 2:
            i, j = ids
            v = p[i] - p[j]
            d = np.linalg.norm(v)
            ws = w[i] + w[j]
            if ws == 0 or d == 0:
                continue
            c = d - float(oc.rest[prim])
            acc[i] += -w[i] / ws * c * v / d * kd
            acc[j] += w[j] / ws * c * v / d * kd
        else:
            acc[ids] += bend_fn(prim, ids, p.astype(np.float32)) - p[ids]
    out = p.copy()
    free = (w != 0) & (cnt > 0)
    out[free] += (omega / cnt[free])[:, None] * acc[free]
    return out


@pytest.mark.parametrize("omega", [1.0, 1.5])
def test_distance_sweep_matches_float64_restatement(orc, omega):
    oc, p = stretched_grid(orc)
    for sweep in (0, 3):
        got = oc.jacobi_sweep(sweep, omega, p)
        ref = numpy_jacobi(oc, p, sweep, omega)
        assert np.abs(got - ref).max() < 2e-6, np.abs(got - ref).max()


def test_bending_sweep_averages_single_prim_projections(orc):
    oc, p = stretched_grid(orc, bending=True)
    kb = float(oc.kb[0])
    ids4 = (orc.C.c_int * 4)(0, 1, 2, 3)

    def bend(prim, ids, p32):
        loc = np.ascontiguousarray(p32[ids], np.float32)
        wl = np.ascontiguousarray(oc.w[ids], np.float32)
        oc.L.orc_project_bending(ids4, float(oc.rest[prim]), kb, orc.F(wl), orc.F(loc))
        return loc.astype(np.float64)

    oc.L.orc_project_bending.argtypes = [orc.C.POINTER(orc.C.c_int), orc.C.c_float, orc.C.c_float, orc._fp, orc._fp]
    oc.L.orc_project_bending.restype = None
    got = oc.jacobi_sweep(0, 1.3, p)
    ref = numpy_jacobi(oc, p, 0, 1.3, bend_fn=bend)
    assert np.abs(got - ref).max() < 2e-6
    assert len(oc.prim_list) > 2 * 6 * 7  # bending prims present


def test_single_prim_is_satisfied(orc):
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]], np.float32)
    oc = orc.OracleCloth(pos, np.zeros((0, 3), np.int32), np.array([[0, 1]], np.int32), iterations=1)
    p = np.array([[0.1, 0.2, -0.1], [1.4, -0.3, 0.2]], np.float32)
    out = oc.jacobi_sweep(0, 1.0, p)
    assert abs(np.linalg.norm(out[0].astype(np.float64) - out[1]) - 1.0) < 1e-6
    # equal weights: both ends move by the same amount in opposite directions
    assert np.allclose(out[0] - p[0], -(out[1] - p[1]), atol=1e-7)


def test_rest_state_and_pins_do_not_move(orc):
    oc, p = stretched_grid(orc, bending=True)
    rest = oc.x.copy()
    assert np.array_equal(oc.jacobi_sweep(0, 1.5, rest), rest)
    out = oc.jacobi_sweep(0, 1.5, p)
    assert np.array_equal(out[[0, 6]], p[[0, 6]])


def test_vertex_numbering_does_not_matter(orc):
    oc, p = stretched_grid(orc, seed=2)
    nv = len(p)
    perm = np.random.default_rng(5).permutation(nv)  # new index of old vertex i is inv[i]
    inv = np.empty(nv, np.int64)
    inv[perm] = np.arange(nv)
    pairs = np.array([oc.indices[st:st + 2] for st, n in oc.prim_list], np.int64)
    oc2 = orc.OracleCloth(oc.x[perm], np.zeros((0, 3), np.int32), inv[pairs].astype(np.int32),
                          pins=[int(inv[0]), int(inv[6])], stiffness=(0.8, 0.5, 1.0))
    a = oc.jacobi_sweep(1, 1.5, p)
    b = oc2.jacobi_sweep(1, 1.5, p[perm])
    assert np.abs(a[perm] - b).max() < 1e-6


def test_residual_falls_over_sweeps(orc):
    # stiffness 1: k = 1 every sweep (the reference's schedule for s < 1 shrinks k with the sweep index)
    oc, p = stretched_grid(orc, rx=16, ry=12, seed=1, stiffness=(1.0, 1.0, 1.0))
    r0 = oc.residual(p)[0]
    q = p
    rs = []
    for s in range(60):
        q = oc.jacobi_sweep(s, 1.5, q)
        rs.append(oc.residual(q)[0])
    assert rs[-1] < 0.2 * r0
    assert rs[-1] < rs[9] < rs[0]
