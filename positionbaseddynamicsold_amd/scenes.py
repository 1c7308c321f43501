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

from dataclasses import dataclass, field

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

    cid = ctx.add_grid(spec.rx, spec.ry, spec.sx, spec.sy, jitter=spec.jitter, seed=spec.seed)
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
