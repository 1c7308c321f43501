"""Synthetic workloads of the benchmark configurations, generated from a
seeded RNG so the GPU library and the CPU oracle see identical inputs.

config 1: 32x32 cloth (spacing 1/31 m), two top corners pinned, jitter +-1e-4 m,
          g = -9.8, dt = 1/600 s (1/60 s frame / 10 substeps), 20 sweeps, s = 1.
config 5: 512 cloths of uneven size -- side U{16..256} vertices, aspect
          U(0.5, 2) -- each 1 m wide and `aspect` m deep, 0-4 random boundary
          pins, its own sphere collider of radius U(0.2, 0.4) m centred under
          the cloth centre just below the cloth plane; packed into one context.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np

DT = 1.0 / 600.0


@dataclass
class ClothSpec:
    rx: int
    ry: int
    sx: float
    sy: float
    pins: list = field(default_factory=list)
    jitter: float = 0.0
    seed: int = 0
    sphere: tuple | None = None          # ((cx, cy, cz), r)
    wind: tuple = (0.0, 0.0, 0.0)
    wind_drag: float = 0.0
    origin: tuple = (0.0, 0.0, 0.0)      # translation after generation

    @property
    def nv(self) -> int:
        return self.rx * self.ry

    @property
    def ne(self) -> int:
        rx, ry = self.rx, self.ry
        return (ry - 1) * (3 * rx - 2) + (rx - 1)


def config1(seed: int = 0) -> ClothSpec:
    return ClothSpec(32, 32, 1.0, 1.0, pins=[0, 31], jitter=1e-4, seed=seed)


def boundary_vertices(rx: int, ry: int) -> np.ndarray:
    b = set(range(rx)) | set(range((ry - 1) * rx, ry * rx))
    b |= {i * rx for i in range(ry)} | {i * rx + rx - 1 for i in range(ry)}
    return np.array(sorted(b), np.int64)


def config5(seed: int = 0, count: int = 512) -> list[ClothSpec]:
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        side = int(rng.integers(16, 257))
        aspect = float(rng.uniform(0.5, 2.0))
        rx, ry = side, max(2, int(round(side * aspect)))
        sx, sy = 1.0, float(np.float32(aspect))
        npins = int(rng.integers(0, 5))
        pins = sorted(rng.choice(boundary_vertices(rx, ry), npins, replace=False).tolist()) if npins else []
        r = float(np.float32(rng.uniform(0.2, 0.4)))
        centre = (float(np.float32(0.5 * sx)), float(np.float32(-r - 0.02)), float(np.float32(0.5 * sy)))
        specs.append(ClothSpec(rx, ry, sx, sy, pins=pins, jitter=1e-4, seed=seed * 100003 + i,
                               sphere=(centre, r)))
    return specs


def add_to_context(ctx, spec: ClothSpec, bending: bool = False, stiffness=(1.0, 1.0, 1.0)) -> int:
    from . import capi

    cid = ctx.add_grid(spec.rx, spec.ry, spec.sx, spec.sy, origin=spec.origin, jitter=spec.jitter, seed=spec.seed)
    ctx.set_constraint_option(cid, capi.DISTANCE | (capi.BENDING if bending else 0), stiffness)
    if spec.pins:
        ctx.set_pins(cid, spec.pins)
    kw = {}
    if spec.sphere is not None:
        kw.update(sphere_on=1, sphere_c=spec.sphere[0], sphere_r=spec.sphere[1])
    if spec.wind_drag:
        kw.update(wind=spec.wind, wind_drag=spec.wind_drag)
    if kw:
        ctx.set_params(cid, **kw)
    return cid


def icosphere(subdiv: int = 4, radius: float = 1.0, centre=(0.0, 0.0, 0.0)):
    """Icosahedron subdivided `subdiv` times, projected to the sphere; faces
    wound outward.  subdiv 4 gives 2562 vertices / 5120 faces."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    pos = (np.array(verts) * radius + np.asarray(centre)).astype(np.float32)
    return pos, np.array(f, np.int32)


def config4(seed: int = 0) -> ClothSpec:
    """Config-1 cloth with a random wind (speed U(0,5) m/s, horizontal
    direction uniform, drag 1/s), 0-4 random boundary pins and a sphere
    (r 0.3 m) at a random position under the cloth."""
    rng = np.random.default_rng(10_000 + seed)
    s = config1(seed)
    npins = int(rng.integers(0, 5))
    s.pins = sorted(rng.choice(boundary_vertices(32, 32), npins, replace=False).tolist()) if npins else []
    ang = float(rng.uniform(0, 2 * np.pi))
    spd = float(rng.uniform(0, 5))
    s.wind = (float(np.float32(spd * np.cos(ang))), 0.0, float(np.float32(spd * np.sin(ang))))
    s.wind_drag = 1.0
    c = (float(np.float32(rng.uniform(0.2, 0.8))), float(np.float32(rng.uniform(-0.6, -0.2))),
         float(np.float32(rng.uniform(0.2, 0.8))))
    s.sphere = (c, 0.3)
    return s


# ---------------------------------------------------------------------------
# Scenes with a static triangle mesh and vertex-face CCD
# ---------------------------------------------------------------------------
@dataclass
class CcdScene:
    """A grid cloth, a static triangle mesh and the CCD settings of a config."""
    cloth: ClothSpec
    static_pos: np.ndarray
    static_tris: np.ndarray
    bending: bool = False
    stiffness: tuple = (1.0, 1.0, 1.0)
    iterations: int = 20
    table_bits: int = 20
    cell_scale: float = 1.5       # hash cell = cell_scale x mean distance-prim rest length
    thickness: float = 0.001
    resolve_passes: int = 2
    detect_every: int = 10        # projectConstraintWithColliCPU detects at ii % 10 == 0 and the last sweep
    fold: tuple | None = None     # (angle, radius) of the initial fold, rest state flat
    # degenerate-face bound of the response on |(p2-p1) x (p3-p1)|^2 in m^4;
    # 0 = the reference's 1e-6, which skips every face with edges below ~3.4 cm
    min_normal2: float = 0.0
    # candidate pairs: 0 = the reference's centroid cells, 1 = swept boxes
    broadphase: int = 0
    edge_edge: int = 0            # 1: edge-edge contacts too

    def robust(self):
        """The same scene with the swept-box broadphase and a degenerate-face
        bound far below any face of the scene (1e-20 m^4), so that contacts
        are found and resolved on faces of every size."""
        return replace(self, broadphase=1, min_normal2=1e-20)


def box_mesh(size: float = 0.4, centre=(0.0, 0.2, 0.0), div: int = 16):
    """Closed cube of edge `size`, every face a div x div grid of quads split
    into two triangles, wound outward."""
    h = size / 2.0
    c = np.asarray(centre, np.float64)
    pos, tris = [], []
    # (normal axis, sign); u, v axes chosen so that u x v = outward normal
    faces = [(0, 1, 1, 2), (0, -1, 2, 1), (1, 1, 2, 0), (1, -1, 0, 2), (2, 1, 0, 1), (2, -1, 1, 0)]
    for ax, sg, ua, va in faces:
        base = len(pos)
        for j in range(div + 1):
            for i in range(div + 1):
                q = np.zeros(3)
                q[ax] = sg * h
                q[ua] = -h + size * i / div
                q[va] = -h + size * j / div
                pos.append(c + q)
        for j in range(div):
            for i in range(div):
                a = base + j * (div + 1) + i
                b, d, e = a + 1, a + div + 1, a + div + 2
                tris += [(a, b, e), (a, e, d)]
    return np.asarray(pos, np.float32), np.asarray(tris, np.int32)


def fold_positions(pos: np.ndarray, angle: float, radius: float) -> np.ndarray:
    """Folds a flat cloth (y constant) once about the line through its XZ
    centre with direction (cos a, 0, sin a): points on the positive side roll
    over a half cylinder of the given radius and lie flat on top, 2 * radius
    above (isometric, so edge lengths keep their rest values)."""
    p = pos.astype(np.float64)
    ctr = 0.5 * (p.min(0) + p.max(0))
    d = np.array([np.cos(angle), 0.0, np.sin(angle)])
    nrm = np.array([-np.sin(angle), 0.0, np.cos(angle)])
    rel = p - ctr
    t = rel @ d
    sd = rel @ nrm
    y = p[:, 1].copy()
    s2 = sd.copy()
    bend = (sd > 0) & (sd <= np.pi * radius)
    th = sd[bend] / radius
    s2[bend] = radius * np.sin(th)
    y[bend] += radius * (1.0 - np.cos(th))
    top = sd > np.pi * radius
    s2[top] = -(sd[top] - np.pi * radius)
    y[top] += 2.0 * radius
    out = ctr + np.outer(t, d) + np.outer(s2, nrm)
    out[:, 1] = y
    return out.astype(np.float32)


def config2(n: int = 256, subdiv: int = 4, seed: int = 0) -> CcdScene:
    """n x n cloth (1 m, jitter 1e-4), distance + bending 0.1, dropped from
    0.5 m above a static icosphere (r 0.3 m at the origin; subdivision 4 =
    2,562 vertices); hashed grid with cell 1.5 x mean edge, table 2^20."""
    cloth = ClothSpec(n, n, 1.0, 1.0, pins=[], jitter=1e-4, seed=seed, origin=(-0.5, 0.5, -0.5))
    sp, st = icosphere(subdiv, 0.3)
    return CcdScene(cloth, sp, st, bending=True, stiffness=(1.0, 0.1, 1.0), iterations=20, table_bits=20)


def config3(n: int = 512, seed: int = 0, box_div: int = 64) -> CcdScene:
    """n x n cloth (1 m), folded once about a random axis through its centre
    (fold radius 2.5 grid spacings: 4.9 mm, layers 1 cm apart at n = 512),
    5 mm above a static 0.4 m box mesh; 0-4 random boundary pins; self and
    cloth-box contact, table 2^22, 30 sweeps."""
    rng = np.random.default_rng(30_000 + seed)
    npins = int(rng.integers(0, 5))
    pins = sorted(rng.choice(boundary_vertices(n, n), npins, replace=False).tolist()) if npins else []
    angle = float(rng.uniform(0.0, np.pi))
    cloth = ClothSpec(n, n, 1.0, 1.0, pins=pins, jitter=1e-4, seed=seed, origin=(-0.5, 0.405, -0.5))
    bp, bt = box_mesh(0.4, (0.0, 0.2, 0.0), box_div)
    return CcdScene(cloth, bp, bt, iterations=30, table_bits=22, fold=(angle, 2.5 / (n - 1)))


def hash_cell(rest: np.ndarray, scale: float) -> float:
    """scale x mean rest length, with the library's arithmetic (double sum in
    prim order, float result)."""
    s = np.cumsum(rest.astype(np.float64))[-1] if len(rest) else 0.0
    return float(np.float32(np.float64(np.float32(scale)) * (s / len(rest))))


def add_ccd_scene(ctx, sc: CcdScene, graph: int = 0):
    """Adds the cloth and the static mesh, initialises, applies the fold and
    the collision settings; returns (cloth id, static id, hash cell)."""
    from . import capi

    c = sc.cloth
    cid = ctx.add_grid(c.rx, c.ry, c.sx, c.sy, origin=c.origin, jitter=c.jitter, seed=c.seed)
    ctx.set_constraint_option(cid, capi.DISTANCE | (capi.BENDING if sc.bending else 0), sc.stiffness)
    if c.pins:
        ctx.set_pins(cid, c.pins)
    sid = ctx.add_mesh(sc.static_pos, sc.static_tris)
    ctx.set_constraint_option(sid, 0)
    ctx.set_pins(sid, np.arange(len(sc.static_pos)))
    ctx.init()
    _, _, rest = ctx.prims(cid)
    cell = hash_cell(rest[: ctx.info(cid).ne], sc.cell_scale)
    if sc.fold is not None:
        ctx.set(cid, capi.STATE_X, fold_positions(ctx.get(cid), *sc.fold))
    ctx.set_solver(dt=DT, iterations=sc.iterations, graph=graph)
    ctx.set_collision(1, capi.hash_desc(1, cell=(cell, cell, cell), table_bits=sc.table_bits), cell_scale=0.0,
                      thickness=sc.thickness, resolve_passes=sc.resolve_passes, detect_every=sc.detect_every,
                      min_normal2=sc.min_normal2, broadphase=sc.broadphase, edge_edge=sc.edge_edge)
    return cid, sid, cell
