"""Pins of the oracle's edge-edge CCD (orc_ee_ccd, orc_ee_fix,
orc_ee_candidates, orc_ccd_resolve_ee).  The reference declares
Contact::EE (CCD_Basic.h:14) and leaves it unimplemented, so these check the
mathematics rather than reference vectors:

- closed-form crossings: time of impact, closest-point weights, normal;
- misses: edges that pass beside each other, edges already far apart;
- the response: a static edge leaves the whole gap to the other edge, which
  ends exactly 2 x thickness away; two free edges move symmetrically;
- edges side by side in one sheet (no separation along their normal at the
  start) are neither contacts nor pushed apart;
- the candidate set equals an independent numpy brute force over triangle
  pairs whose centroid cells are at most one cell apart, and never pairs
  edges that share a vertex or whose vertices are mesh neighbours.
"""
import numpy as np

F32 = np.float32


def seg(a, b):
    return [np.array(a, F32), np.array(b, F32)]


def test_crossing_edges_time_weights_normal(orc):
    # a: static along x; b: along z at y = 0.1 moving to y = -0.1 (crosses at t = 0.5)
    x = seg((-1, 0, 0), (1, 0, 0)) + seg((0, 0.1, -1), (0, 0.1, 1))
    p = seg((-1, 0, 0), (1, 0, 0)) + seg((0, -0.1, -1), (0, -0.1, 1))
    hit, t, n, w = orc.ee_ccd(x, p, 0.001)
    assert hit
    assert abs(t - 0.5) < 1e-6
    assert np.allclose(w, [0.5, 0.5, 0.5, 0.5], atol=1e-6)
    assert np.allclose(n, [0.0, -1.0, 0.0], atol=1e-7)  # (p1 - p0) x (p3 - p2) = (2,0,0) x (0,0,2)


def test_crossing_off_centre(orc):
    # b crosses a at a = 0.25 along a, 0.75 along b, at t = 0.2
    x = seg((0, 0, 0), (4, 0, 0)) + seg((1, 0.2, -3), (1, 0.2, 1))
    p = seg((0, 0, 0), (4, 0, 0)) + seg((1, -0.8, -3), (1, -0.8, 1))
    hit, t, n, w = orc.ee_ccd(x, p, 0.0005)
    assert hit and abs(t - 0.2) < 1e-6
    assert np.allclose(w, [0.75, 0.25, 0.25, 0.75], atol=1e-6)


def test_passing_beside_and_far_apart_are_misses(orc):
    # b moves through the plane of a, but beyond a's end (x = 2 > 1)
    x = seg((-1, 0, 0), (1, 0, 0)) + seg((2, 0.1, -1), (2, 0.1, 1))
    p = seg((-1, 0, 0), (1, 0, 0)) + seg((2, -0.1, -1), (2, -0.1, 1))
    assert not orc.ee_ccd(x, p, 0.001)[0]
    # both static and 0.5 apart
    x = seg((-1, 0, 0), (1, 0, 0)) + seg((0, 0.5, -1), (0, 0.5, 1))
    assert not orc.ee_ccd(x, x, 0.001)[0]


def test_proximity_at_the_end_of_the_step(orc):
    # no crossing, but b ends within 2 x thickness of a
    x = seg((-1, 0, 0), (1, 0, 0)) + seg((0, 0.1, -1), (0, 0.1, 1))
    p = seg((-1, 0, 0), (1, 0, 0)) + seg((0, 0.003, -1), (0, 0.003, 1))
    hit, t, _, _ = orc.ee_ccd(x, p, 0.002)
    assert hit and t == 1.0
    assert not orc.ee_ccd(x, p, 0.001)[0]


def test_response_static_edge_gives_the_whole_gap(orc):
    th = 0.005
    X = seg((-1, 0, 0), (1, 0, 0)) + seg((0, 0.1, -1), (0, 0.1, 1))
    q = seg((-1, 0, 0), (1, 0, 0)) + seg((0, 0.001, -1), (0, 0.001, 1))
    d = orc.ee_fix(X, q, th)
    assert d is not None
    # free: the closest point of b moves up by half the gap, a down by half
    assert np.allclose(d[0], d[1]) and np.allclose(d[2], d[3])
    assert np.allclose(d[:2, 1], -(2 * th - 0.001) / 2, atol=1e-8)
    assert np.allclose(d[2:, 1], (2 * th - 0.001) / 2, atol=1e-8)
    # with a static (w = 0), b takes the whole gap and ends 2 x thickness away
    x = np.array(X, F32)
    p = np.array(q, F32)
    w = np.array([0, 0, 1, 1], F32)
    out = orc.ccd_resolve_ee(None, 0, np.array([[0, 1, 2, 3]], np.int32), np.zeros((0, 3), np.int32), x, p, w, th, 1)
    assert np.array_equal(out[:2], p[:2])
    assert np.allclose(out[2:, 1], 2 * th, atol=1e-7)
    # separated by 2 x thickness: nothing left to fix
    assert orc.ee_fix(X, [out[0], out[1], out[2], out[3]], th) is None


def test_response_is_oriented_by_the_start_side(orc):
    # b started below a and ends just above it: pushed back below
    th = 0.005
    X = seg((-1, 0, 0), (1, 0, 0)) + seg((0, -0.1, -1), (0, -0.1, 1))
    q = seg((-1, 0, 0), (1, 0, 0)) + seg((0, 0.002, -1), (0, 0.002, 1))
    d = orc.ee_fix(X, q, th)
    assert d is not None and (d[2:, 1] < 0).all() and (d[:2, 1] > 0).all()


def test_side_by_side_edges_are_not_contacts(orc):
    # two edges of one flat sheet, 1 mm apart in its plane: within the contact
    # gap, but with no separation along their normal to restore
    x = seg((0, 0, 0), (1, 0, 0)) + seg((0.5, 0, 0.001), (0.5, 0, 1))
    p = [v + np.array([0, -0.01, 0], F32) for v in x]
    assert not orc.ee_ccd(x, p, 0.002)[0]
    assert orc.ee_fix(x, p, 0.002) is None
    # the same edges starting 3 mm apart across the sheet normal are
    x2 = seg((0, 0, 0), (1, 0, 0)) + seg((0.5, 0.003, 0.001), (0.5, 0.003, 1))
    assert orc.ee_ccd(x2, x2, 0.002)[0]


def test_parallel_edges_are_skipped_by_the_response(orc):
    X = seg((-1, 0, 0), (1, 0, 0)) + seg((-1, 0.1, 0), (1, 0.1, 0))
    q = seg((-1, 0, 0), (1, 0, 0)) + seg((-1, 0.001, 0), (1, 0.001, 0))
    assert orc.ee_fix(X, q, 0.005) is None


def folded(orc, n=14, seed=4):
    pos, tris = orc.grid_mesh(n, n, 1.0, 1.0, jitter=1e-4, seed=seed)
    pos = pos.copy()
    rng = np.random.default_rng(seed)
    pos[:, 1] = (0.06 * np.sin(9 * pos[:, 0]) + 0.02 * rng.standard_normal(len(pos))).astype(F32)
    return pos.astype(F32), tris


def test_candidates_equal_numpy_brute_force(orc):
    pos, tris = folded(orc)
    cell = 0.12
    g = orc.grid(1, cell=(cell, cell, cell), table_bits=10)
    cand, edges = orc.ee_candidates(g, pos, tris)
    assert len(cand) > 100
    # independent: cells from the float32 centroid, edges from the triangles
    cen = ((pos[tris[:, 0]] + pos[tris[:, 1]]) + pos[tris[:, 2]]) / F32(3)
    cells = np.floor((cen - np.array(g.start, F32)) / np.array(g.cell, F32)).astype(np.int64)
    edge_id = {tuple(e): i for i, e in enumerate(edges.tolist())}
    nbrs = {}
    for t3 in tris.tolist():
        for k in range(3):
            a, b = t3[k], t3[(k + 1) % 3]
            nbrs.setdefault(a, set()).add(b)
            nbrs.setdefault(b, set()).add(a)
    ref = set()
    for t in range(len(tris)):
        for f in range(t + 1, len(tris)):
            if np.abs(cells[t] - cells[f]).max() > 1:
                continue
            for ka in range(3):
                a = tuple(sorted((int(tris[t][ka]), int(tris[t][(ka + 1) % 3]))))
                for kb in range(3):
                    b = tuple(sorted((int(tris[f][kb]), int(tris[f][(kb + 1) % 3]))))
                    if set(a) & set(b) or (nbrs[a[0]] | nbrs[a[1]]) & set(b):
                        continue
                    ea, eb = edge_id[a], edge_id[b]
                    ref.add((min(ea, eb), max(ea, eb)))
    assert set(map(tuple, cand.tolist())) == ref
    assert (np.diff(cand[:, 0]) >= 0).all()
    for ea, eb in cand[:200]:
        assert not set(edges[ea]) & set(edges[eb])


def test_edge_table_first_sighting(orc):
    tris = np.array([[0, 1, 2], [2, 1, 3]], np.int32)
    tedge, edges = orc.edge_table(tris)
    assert edges.tolist() == [[0, 1], [1, 2], [0, 2], [1, 3], [2, 3]]
    assert tedge.tolist() == [0, 1, 2, 1, 3, 4]
