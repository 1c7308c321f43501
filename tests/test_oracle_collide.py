"""Pins the oracle's broadphase and CCD against closed forms, brute force and
numpy root finding.  CPU only."""
import itertools

import numpy as np
import pytest


def tri_at(centroids):
    """One triangle per centroid: vertices centroid + offsets summing to 0."""
    off = np.array([[0.01, 0.0, 0.0], [-0.005, 0.01, 0.0], [-0.005, -0.01, 0.0]], np.float32)
    pos = (np.repeat(np.asarray(centroids, np.float32), 3, 0) + np.tile(off, (len(centroids), 1))).astype(np.float32)
    tris = np.arange(3 * len(centroids), dtype=np.int32).reshape(-1, 3)
    return pos, tris


def test_dense_keys_known_cells(orc):
    # grid 4x3x2 cells of size 1 centred at (2, 1.5, 1): start = (0, 0, 0)
    g = orc.grid(0, centre=(2.0, 1.5, 1.0), size=(4, 3, 2), cell=(1.0, 1.0, 1.0))
    assert list(g.start) == [0.0, 0.0, 0.0]
    cents = [(0.5, 0.5, 0.5), (3.5, 2.5, 1.5), (2.25, 0.75, 1.25), (-0.5, 0.5, 0.5), (0.5, 3.5, 0.5), (4.5, 0.5, 0.5)]
    pos, tris = tri_at(cents)
    keys, ids = orc.tri_keys(g, pos, tris)
    # key = (z*gy + y)*gx + x, -1 outside (calcGridHashCPU)
    assert keys.tolist() == [0, (1 * 3 + 2) * 4 + 3, (1 * 3 + 0) * 4 + 2, -1, -1, -1]
    assert ids.tolist() == list(range(6))


def test_stable_sort_and_ranges_brute_force(orc):
    rng = np.random.default_rng(0)
    keys = rng.integers(-1, 20, 500).astype(np.int32)
    ids = np.arange(500, dtype=np.int32)
    sk, si = orc.sorted_keys(keys, ids)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(si, order) and np.array_equal(sk, keys[order])
    g = orc.grid(0, size=(20, 1, 1))
    cs, ce = orc.cell_ranges(g, sk)
    for k in range(20):
        where = np.nonzero(sk == k)[0]
        if len(where):
            assert (cs[k], ce[k]) == (where.min(), where.max())
        else:
            assert (cs[k], ce[k]) == (-1, -1)


def brute_pairs(keyfun_cell, pos, tris):
    # (v, f): f's centroid cell within one cell of t's on every axis, both
    # triangles with a valid key (inside the grid in dense mode)
    cen = pos[tris].astype(np.float32).sum(1, dtype=np.float32) / np.float32(3)
    out = set()
    cells = [keyfun_cell(c) for c in cen]
    for t in range(len(tris)):
        gp, key = cells[t]
        if key < 0:
            continue
        for f in range(len(tris)):
            gf, kf = cells[f]
            if f == t or kf < 0 or any(abs(gf[k] - gp[k]) > 1 for k in range(3)):
                continue
            for v in tris[t]:
                if v not in tris[f]:
                    out.add((int(v) << 32) | f)
    return np.array(sorted(out), np.int64)


@pytest.mark.parametrize("mode", [0, 1])
def test_candidate_pairs_brute_force(orc, mode):
    rng = np.random.default_rng(mode)
    nv, nt = 90, 70
    pos = rng.uniform(-1, 1, (nv, 3)).astype(np.float32)
    tris = np.array([rng.choice(nv, 3, replace=False) for _ in range(nt)], np.int32)
    g = orc.grid(mode, centre=(0, 0, 0), size=(5, 5, 5), cell=(0.35, 0.35, 0.35), table_bits=5)
    start = np.array(g.start, np.float32)
    cell = np.array(g.cell, np.float32)

    def key_of(gp):
        if mode == 0:
            if any(c < 0 or c >= 5 for c in gp):
                return -1
            return (gp[2] * 5 + gp[1]) * 5 + gp[0]
        h = (np.uint32(gp[0] & 0xffffffff) * np.uint32(73856093)) ^ (np.uint32(gp[1] & 0xffffffff) * np.uint32(19349663)) ^ \
            (np.uint32(gp[2] & 0xffffffff) * np.uint32(83492791))
        return int(h & np.uint32(31))

    def cell_of(c):
        gp = tuple(int(np.floor(np.float32((c[k] - start[k])) / cell[k])) for k in range(3))
        return gp, key_of(gp)

    with np.errstate(over="ignore"):
        expect = brute_pairs(cell_of, pos, tris)
    got = orc.candidate_pairs(g, pos, tris)
    assert np.array_equal(got, expect)
    assert len(got) > 0


def test_hashed_pairs_independent_of_table_size(orc):
    # a 5-bit table collides many cells into one slot; the set must not change
    # with the table size and must equal the dense grid's when every triangle
    # lies inside the dense grid
    rng = np.random.default_rng(11)
    nv, nt = 160, 140
    pos = rng.uniform(-0.8, 0.8, (nv, 3)).astype(np.float32)
    tris = np.array([rng.choice(nv, 3, replace=False) for _ in range(nt)], np.int32)
    cell, size = (0.3, 0.3, 0.3), (8, 8, 8)  # dense grid spans [-1.2, 1.2]^3
    sets = [orc.candidate_pairs(orc.grid(1, size=size, cell=cell, table_bits=b), pos, tris) for b in (3, 5, 8, 16)]
    dense = orc.candidate_pairs(orc.grid(0, size=size, cell=cell), pos, tris)
    assert len(dense) > 0
    for s_ in sets:
        assert np.array_equal(s_, dense)


def test_cubic_roots_against_numpy(orc):
    import ctypes as C
    L = orc.lib()
    rng = np.random.default_rng(2)
    x = np.zeros(3)
    for _ in range(300):
        r = np.sort(rng.uniform(-2, 2, 3))
        a = rng.uniform(0.5, 2.0) * rng.choice([-1, 1])
        c = a * np.poly(r)  # a x^3 + b x^2 + c x + d with roots r
        n = L.orc_solve_cubic(c[0], c[1], c[2], c[3], x.ctypes.data_as(C.POINTER(C.c_double)))
        got = np.sort(x[:n])
        if np.min(np.diff(r)) > 1e-2:
            assert n == 3
            assert np.allclose(got, r, atol=1e-5)
    # one real root: the other two complex
    c = np.poly([0.3, 1 + 2j, 1 - 2j]).real
    n = L.orc_solve_cubic(c[0], c[1], c[2], c[3], x.ctypes.data_as(C.POINTER(C.c_double)))
    assert n == 1 and abs(x[0] - 0.3) < 1e-6


def moving_vertex_scene(xy, z0, z1):
    # static face 0 = (1,2,3) in the z = 0 plane; vertex 0 belongs to a small
    # vertical triangle 1 = (0,4,5) that translates with it (candidates come
    # from triangle vertices only, CCD.cpp:1541-1556)
    tris = np.array([[1, 2, 3], [0, 4, 5]], np.int32)
    x = np.array([[xy[0], xy[1], z0], [0, 0, 0], [1, 0, 0], [0, 1, 0],
                  [xy[0] + 0.05, xy[1], z0 + 0.1], [xy[0], xy[1] + 0.05, z0 + 0.1]], np.float32)
    p = x.copy()
    p[[0, 4, 5], 2] += np.float32(z1 - z0)
    return tris, x, p


def test_vf_ccd_constructed_trajectory(orc):
    # vertex from z = +1 to z = -1 through an interior point: coplanar at t = 0.5
    tris, x, p = moving_vertex_scene((0.2, 0.2), 1.0, -1.0)
    g = orc.grid(1, cell=(2.0, 2.0, 2.0), table_bits=8)
    arr, n = orc.ccd_detect(g, x, p, tris, 0.01)
    vf, t, nn, w = orc.contacts_arrays(arr, n)
    hit = [i for i in range(n) if vf[i].tolist() == [0, 0]]
    assert len(hit) == 1
    i = hit[0]
    assert abs(t[i] - 0.5) < 1e-6
    assert np.allclose(np.abs(nn[i]), [0, 0, 1])
    assert np.allclose(w[i, 1:].sum(), 1.0, atol=1e-6)
    # outside the face: no vertex-0 / face-0 contact
    tris, x2, p2 = moving_vertex_scene((0.8, 0.8), 1.0, -1.0)
    arr, n = orc.ccd_detect(g, x2, p2, tris, 0.01)
    vf, *_ = orc.contacts_arrays(arr, n)
    assert [0, 0] not in vf.tolist()


def test_resolve_separates_to_thickness(orc):
    # vertex ended 0.001 below a face it started above; after resolve the
    # vertex is back on the collision-free side at >= 2h - tol
    h = 0.01
    tris, x, p = moving_vertex_scene((0.25, 0.25), 0.05, -0.001)
    g = orc.grid(1, cell=(2.0, 2.0, 2.0), table_bits=8)
    arr, n = orc.ccd_detect(g, x, p, tris, h)
    vf, *_ = orc.contacts_arrays(arr, n)
    assert [0, 0] in vf.tolist()
    # keep only the vertex-0 / face-0 contact
    keep = [i for i in range(n) if vf[i].tolist() == [0, 0]]
    arr = (type(arr[0]) * 1)(arr[keep[0]])
    n = 1
    for w, tol in ((np.ones(6, np.float32), 1e-5), (np.array([1, 0, 0, 0, 1, 1], np.float32), 1e-5)):
        q = orc.ccd_resolve(arr, n, tris, x, p, w, h, 8)
        nrm = np.cross(q[2] - q[1], q[3] - q[1])
        nrm /= np.linalg.norm(nrm)
        d = float(np.dot(q[0] - q[1], nrm))
        assert d >= 2 * h - tol
        if w[1] == 0:
            assert np.array_equal(q[1:4], p[1:4])  # static face untouched


def test_crossings_closed_forms(orc):
    # the reference's post-resolve check (NarrowVFCCDTest): a vertex through
    # the interior of a static face crosses it at t = 1/2 (so do vertices 4
    # and 5, which move with it over interior points); one stopping 1 mm short
    # of it, or passing beyond the hypotenuse, does not
    tris, x, p = moving_vertex_scene((0.2, 0.2), 1.0, -1.0)
    assert orc.crossings(x, p, tris, 0.01).tolist() == [[0, 0], [4, 0], [5, 0]]
    tris, x, p = moving_vertex_scene((0.2, 0.2), 1.0, 0.001)
    assert len(orc.crossings(x, p, tris, 0.01)) == 0
    tris, x, p = moving_vertex_scene((0.8, 0.8), 1.0, -1.0)
    assert [0, 0] not in orc.crossings(x, p, tris, 0.01).tolist()
    # a face swept past a static vertex crosses it too (vertex 0 is the static one)
    tris, x, p = moving_vertex_scene((0.2, 0.2), 0.0, 0.0)
    x[1:4, 2] = -1.0
    p[1:4, 2] = 1.0
    assert [0, 0] in orc.crossings(x, p, tris, 0.01).tolist()
    # the response leaves a crossing vertex on its start side, 2h off the face
    h = 0.01
    tris, x, p = moving_vertex_scene((0.25, 0.25), 0.05, -0.001)
    assert orc.crossings(x, p, tris, h).tolist() == [[0, 0]]
    g = orc.grid(1, cell=(2.0, 2.0, 2.0), table_bits=8)
    arr, n = orc.ccd_detect(g, x, p, tris, h)
    q = orc.ccd_resolve(arr, n, tris, x, p, np.array([1, 0, 0, 0, 1, 1], np.float32), h, 2)
    assert len(orc.crossings(x, q, tris, h)) == 0


def contacts(orc, vf):
    arr = (orc.OrcContact * len(vf))()
    for i, (v, f) in enumerate(vf):
        arr[i].v, arr[i].f = v, f
    return arr


def test_resolve_averages_two_contacts_of_one_vertex(orc):
    # One free vertex that ended below two static faces: the plane y = 0
    # (normal +y) and the plane x = 0 (normal +x).  It started above both.
    # Each contact moves it to 2h off its plane, the whole way since the face
    # is static (2 x the half fix); the vertex takes the mean of the two fixes:
    #   h = 1/8, p = (-1/2, -1/4, 0): fixes (0, 2h + 1/4, 0) = (0, 1/2, 0) and
    #   (2h + 1/2, 0, 0) = (3/4, 0, 0); mean (3/8, 1/4, 0) -> (-1/8, 0, 0).
    h = 0.125
    face_y = [(-1, 0, -1), (-1, 0, 1), (1, 0, -1)]    # (e1 x e2) = (0, 4, 0)
    face_x = [(0, -1, -1), (0, 1, -1), (0, -1, 1)]    # (e1 x e2) = (4, 0, 0)
    x = np.array(face_y + face_x + [(0.5, 0.5, 0.0)], np.float32)
    p = x.copy()
    p[6] = (-0.5, -0.25, 0.0)
    tris = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    w = np.array([0, 0, 0, 0, 0, 0, 1], np.float32)
    q = orc.ccd_resolve(contacts(orc, [(6, 0), (6, 1)]), 2, tris, x, p, w, h, 1)
    assert q[6].tolist() == [-0.125, 0.0, 0.0]
    assert np.array_equal(q[:6], p[:6])
    # each contact alone moves the vertex by its own fix
    q0 = orc.ccd_resolve(contacts(orc, [(6, 0)]), 1, tris, x, p, w, h, 1)
    q1 = orc.ccd_resolve(contacts(orc, [(6, 1)]), 1, tris, x, p, w, h, 1)
    assert q0[6].tolist() == [-0.5, 0.25, 0.0] and q1[6].tolist() == [0.25, -0.25, 0.0]


def test_resolve_averages_face_fixes_of_two_contacts(orc):
    # Two free vertices below one free face (normal +y, all weights 1), both
    # started above it.  Each contact splits its fix: the vertex moves by half
    # the gap to 2h, the face vertices by the other half along -n, spread by
    # the vertex's barycentric weights b_k as b_k / sum(b^2).  A face vertex
    # that both contacts move takes the mean of the two.
    #   h = 1/8; face (0,0,0), (0,0,1), (1,0,0)
    #   v1 at (1/4, -1/8, 1/4): b = (1/2, 1/4, 1/4), sum b^2 = 3/8,
    #      half gap (2h + 1/8) / 2 = 3/16: v1 +3/16; face -1/4, -1/8, -1/8
    #   v2 at (0, -3/8, 1/2): b = (1/2, 1/2, 0), sum b^2 = 1/2,
    #      half gap (2h + 3/8) / 2 = 5/16: v2 +5/16; face -5/16, -5/16, 0
    #   means: face -9/32, -7/32, -1/16
    h = 0.125
    x = np.array([(0, 0, 0), (0, 0, 1), (1, 0, 0), (0.25, 0.125, 0.25), (0.0, 0.125, 0.5)], np.float32)
    p = x.copy()
    p[3, 1] = -0.125
    p[4, 1] = -0.375
    tris = np.array([[0, 1, 2]], np.int32)
    w = np.ones(5, np.float32)
    q = orc.ccd_resolve(contacts(orc, [(3, 0), (4, 0)]), 2, tris, x, p, w, h, 1)
    dy = (q - p)[:, 1].tolist()
    assert dy == [-9 / 32, -7 / 32, -1 / 16, 3 / 16, 5 / 16]
    assert np.array_equal((q - p)[:, [0, 2]], np.zeros((5, 2), np.float32))


def falling_patch_scene(rng, n=7, big=3, drop=0.03):
    """A static plane of `big` x `big` large quads (two triangles each) and an
    n x n jittered patch of small triangles moving down through it."""
    s = np.linspace(-0.5, 0.5, big + 1)
    gx, gz = np.meshgrid(s, s, indexing="ij")
    plane = np.stack([gx.ravel(), np.zeros(gx.size), gz.ravel()], 1)
    ptris = []
    for i in range(big):
        for j in range(big):
            a, b, c, d = i * (big + 1) + j, i * (big + 1) + j + 1, (i + 1) * (big + 1) + j, (i + 1) * (big + 1) + j + 1
            ptris += [[a, c, b], [b, c, d]]
    t = np.linspace(-0.3, 0.3, n)
    hx, hz = np.meshgrid(t, t, indexing="ij")
    patch = np.stack([hx.ravel(), np.full(hx.size, 0.01), hz.ravel()], 1) + rng.uniform(-0.01, 0.01, (n * n, 3))
    qtris = []
    for i in range(n - 1):
        for j in range(n - 1):
            a, b, c, d = i * n + j, i * n + j + 1, (i + 1) * n + j, (i + 1) * n + j + 1
            qtris += [[a, b, c], [b, d, c]]
    x = np.concatenate([plane, patch]).astype(np.float32)
    p = x.copy()
    m = len(plane)
    p[m:] += rng.uniform(-0.005, 0.005, (n * n, 3)).astype(np.float32)
    p[m:, 1] -= np.float32(drop)
    tris = np.concatenate([np.array(ptris), np.array(qtris) + m]).astype(np.int32)
    return x, p, tris


def test_swept_contacts_equal_all_pairs(orc):
    # the swept-box set keeps every pair the narrowphase accepts: contacts
    # over it equal contacts over all (v, f); the set is a strict subset
    rng = np.random.default_rng(21)
    total = 0
    for trial in range(6):
        x, p, tris = falling_patch_scene(rng, n=6 + trial, big=1 + trial % 3, drop=0.005 + 0.01 * trial)
        for th in (1e-4, 2e-3, 1e-2):
            a, n = orc.ccd_detect_swept(x, p, tris, th)
            b, m = orc.ccd_detect_swept(x, p, tris, th, all_pairs=True)
            ca, cb = orc.contacts_arrays(a, n), orc.contacts_arrays(b, m)
            for u, v in zip(ca, cb):
                assert np.array_equal(u, v)
            total += n
            cand = orc.swept_pairs(x, p, tris, th)
            every = orc.swept_pairs(x, p, tris, th, all_pairs=True)
            assert len(cand) < len(every) and np.isin(cand, every).all()
    assert total > 0
    # random soup: skinny triangles, every vertex moving
    for seed in range(4):
        r = np.random.default_rng(100 + seed)
        nv = 90
        x = r.uniform(-0.2, 0.2, (nv, 3)).astype(np.float32)
        p = (x + r.normal(0, 0.05, (nv, 3))).astype(np.float32)
        tris = np.array([r.choice(nv, 3, replace=False) for _ in range(70)], np.int32)
        for th in (1e-3, 2e-2):
            a, n = orc.ccd_detect_swept(x, p, tris, th)
            b, m = orc.ccd_detect_swept(x, p, tris, th, all_pairs=True)
            assert n == m
            for u, v in zip(orc.contacts_arrays(a, n), orc.contacts_arrays(b, m)):
                assert np.array_equal(u, v)


def test_swept_pairs_ignore_vertices_outside_triangles(orc):
    # a vertex in no triangle is never a candidate (candidates come from
    # triangle vertices, CCD.cpp:1541-1556)
    x, p, tris = falling_patch_scene(np.random.default_rng(3), n=5, big=1)
    lone = len(x)
    x = np.concatenate([x, [[0.0, 0.005, 0.0]]]).astype(np.float32)
    p = np.concatenate([p, [[0.0, -0.005, 0.0]]]).astype(np.float32)
    keys = orc.swept_pairs(x, p, tris, 1e-3)
    assert not np.any((keys >> 32) == lone)
    every = orc.swept_pairs(x, p, tris, 1e-3, all_pairs=True)
    assert not np.any((every >> 32) == lone)


def test_resolve_degenerate_bound(orc):
    # VFResolveTest skips a face with |(p2-p1) x (p3-p1)|^2 <= 1e-6: a 1 cm
    # face (|n|^2 = 1e-8) is never resolved with the reference bound and is
    # resolved with a bound below its |n|^2
    h = 1e-3
    tris, x, p = moving_vertex_scene((0.25, 0.25), 0.05, -0.001)
    s = np.float32(0.01)
    x[1:4] *= s
    p[1:4] *= s
    x[[0, 4, 5], :2] *= s
    p[[0, 4, 5], :2] *= s
    arr, n = orc.ccd_detect_swept(x, p, tris, h)
    vf, *_ = orc.contacts_arrays(arr, n)
    keep = [i for i in range(n) if vf[i].tolist() == [0, 0]]
    assert keep
    one = (type(arr[0]) * 1)(arr[keep[0]])
    w = np.array([1, 0, 0, 0, 1, 1], np.float32)
    q = orc.ccd_resolve(one, 1, tris, x, p, w, h, 4)
    assert np.array_equal(q, p)
    q = orc.ccd_resolve(one, 1, tris, x, p, w, h, 4, min_n2=1e-12)
    assert q[0, 2] >= 2 * h - 1e-6
