"""ctypes wrapper of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It never participates in the product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


class OrcParams(C.Structure):
    _fields_ = [("dt", C.c_float), ("gravity", C.c_float * 3), ("damp_factor", C.c_float),
                ("wind", C.c_float * 3), ("wind_drag", C.c_float), ("sphere_on", C.c_int),
                ("sphere_c", C.c_float * 3), ("sphere_r", C.c_float), ("box_on", C.c_int),
                ("box_lo", C.c_float * 3), ("box_hi", C.c_float * 3)]


class OrcCloth(C.Structure):
    _fields_ = [("nv", C.c_int), ("x", _fp), ("v", _fp), ("p", _fp), ("w", _fp), ("np", C.c_int),
                ("prim_list", _ip), ("indices", _ip), ("rest", _fp), ("sorted", _ip), ("ncolors", C.c_int),
                ("worksets", _ip), ("k_dist", _fp), ("k_bend", _fp), ("k_len", C.c_int)]


class OrcGrid(C.Structure):
    _fields_ = [("mode", C.c_int), ("start", C.c_float * 3), ("cell", C.c_float * 3), ("size", C.c_int * 3),
                ("table_bits", C.c_int)]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        sig = {
            "orc_grid_positions": (None, [C.c_int, C.c_int, C.c_float, C.c_float, _fp]),
            "orc_jitter": (None, [_fp, C.c_int, C.c_float, C.c_uint64]),
            "orc_grid_tris": (C.c_int, [C.c_int, C.c_int, _ip]),
            "orc_grid_distance_pairs": (C.c_int, [C.c_int, C.c_int, _ip]),
            "orc_distance_powf": (C.c_float, [_fp, _fp]),
            "orc_distance": (C.c_float, [_fp, _fp]),
            "orc_rest_lengths": (None, [_fp, _ip, C.c_int, _fp]),
            "orc_bending_quads": (C.c_int, [_ip, C.c_int, _ip]),
            "orc_atan2f": (C.c_float, [C.c_float, C.c_float]),
            "orc_dihedral": (C.c_float, [_fp, _fp, _fp, _fp]),
            "orc_dihedral_grad": (None, [_fp, _fp, _fp, _fp, _dp]),
            "orc_dihedral_grad_f64": (None, [_dp, _dp, _dp, _dp, _dp, _dp]),
            "orc_prim_adjacency": (C.c_int, [_ip, C.c_int, _ip, C.c_int, _ip, _ip]),
            "orc_color": (C.c_int, [_ip, C.c_int, _ip, _ip, C.c_int, _ip, _ip]),
            "orc_color_sort": (C.c_int, [_ip, C.c_int, _ip, _ip]),
            "orc_advect": (None, [C.POINTER(OrcParams), C.c_int, _fp, _fp, _fp, _fp]),
            "orc_project_distance": (None, [_ip, C.c_float, C.c_float, _fp, _fp]),
            "orc_project_bending": (None, [_ip, C.c_float, C.c_float, _fp, _fp]),
            "orc_sweep": (None, [C.POINTER(OrcCloth), C.c_int]),
            "orc_colliders": (None, [C.POINTER(OrcParams), C.c_int, _fp, _fp]),
            "orc_integrate": (None, [C.c_int, C.c_float, _fp, _fp, _fp]),
            "orc_residual": (None, [C.POINTER(OrcCloth), _fp, _dp, _dp, _ip]),
            "orc_substep": (C.c_int, [C.POINTER(OrcCloth), C.POINTER(OrcParams), C.c_int, C.c_int, C.c_double, _dp]),
            "orc_substep_omp": (C.c_int, [C.POINTER(OrcCloth), C.POINTER(OrcParams), C.c_int, C.c_int]),
            "orc_substep_jacobi": (C.c_int, [C.POINTER(OrcCloth), C.POINTER(OrcParams), C.c_int, C.c_float, C.c_int,
                                             C.c_double, _dp]),
            "orc_jacobi_sweep": (None, [C.POINTER(OrcCloth), C.c_int, C.c_float, _fp, _fp]),
            "orc_stiffness_table_ref": (None, [C.c_float, C.c_int, _fp]),
            "orc_stiffness_table_const": (None, [C.c_float, C.c_int, C.c_int, _fp]),
            "orc_grid_init": (None, [C.POINTER(OrcGrid), _fp, _ip, _fp]),
            "orc_tri_centroids": (None, [_fp, _ip, C.c_int, _fp]),
            "orc_cell_coords": (None, [C.POINTER(OrcGrid), _fp, _ip]),
            "orc_cell_key": (C.c_int, [C.POINTER(OrcGrid), _ip]),
            "orc_tri_keys": (None, [C.POINTER(OrcGrid), _fp, _ip, C.c_int, _ip, _ip]),
            "orc_stable_sort_pairs": (None, [_ip, _ip, C.c_int]),
            "orc_num_cells": (C.c_int, [C.POINTER(OrcGrid)]),
            "orc_cell_ranges": (None, [_ip, C.c_int, C.c_int, _ip, _ip]),
            "orc_candidate_pairs": (C.c_longlong, [C.POINTER(OrcGrid), _fp, _ip, C.c_int, C.c_int, C.c_longlong, _i64p]),
            "orc_solve_quadratic": (C.c_int, [C.c_double, C.c_double, C.c_double, _dp]),
            "orc_solve_cubic": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, _dp]),
            "orc_barycentric": (None, [_fp, _fp, _fp, _fp, _fp]),
            "orc_vf_ccd": (C.c_int, [_fp] * 8 + [C.c_float, _fp, _fp, _dp]),
            "orc_vf_resolve_test": (C.c_int, [_fp] * 8 + [C.c_float, C.c_float]),
            "orc_vf_resolve": (None, [_fp] * 8 + [C.c_float]),
            "orc_gelu": (None, [C.c_double, _dp, _dp]),
            "orc_mlp_forward": (None, [_dp] * 7 + [C.c_int] + [_dp] * 5),
            "orc_mlp_loss_grad": (C.c_double, [_dp] * 8 + [C.c_int] + [_dp] * 6),
            "orc_adam_step": (None, [_dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]),
            "orc_f32_to_bf16": (C.c_uint16, [C.c_float]),
            "orc_bf16_to_f32": (C.c_float, [C.c_uint16]),
            "orc_splitmix64": (C.c_uint64, [C.POINTER(C.c_uint64)]),
            "orc_uniform_pm1": (C.c_float, [C.POINTER(C.c_uint64)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def F(a):
    return a.ctypes.data_as(_fp)


def I(a):
    return a.ctypes.data_as(_ip)


def D(a):
    return a.ctypes.data_as(_dp)


_libm = C.CDLL("libm.so.6")
_libm.powf.restype = C.c_float
_libm.powf.argtypes = [C.c_float, C.c_float]


def powf(a: float, b: float) -> float:
    return _libm.powf(a, b)


# ---------------------------------------------------------------------------
# Mesh helpers
# ---------------------------------------------------------------------------
def grid_mesh(rx, ry, sx=1.0, sy=1.0, jitter=0.0, seed=0, origin=(0.0, 0.0, 0.0)):
    L = lib()
    pos = np.zeros((rx * ry, 3), np.float32)
    L.orc_grid_positions(rx, ry, sx, sy, F(pos))
    if jitter != 0.0:
        L.orc_jitter(F(pos), rx * ry, jitter, seed)
    if any(o != 0.0 for o in origin):
        pos = (pos + np.asarray(origin, np.float32)).astype(np.float32)
    nt = L.orc_grid_tris(rx, ry, None)
    tris = np.zeros((nt, 3), np.int32)
    L.orc_grid_tris(rx, ry, I(tris))
    return pos, tris


def grid_pairs(rx, ry):
    L = lib()
    ne = L.orc_grid_distance_pairs(rx, ry, None)
    pairs = np.zeros((ne, 2), np.int32)
    L.orc_grid_distance_pairs(rx, ry, I(pairs))
    return pairs


def bending_quads(tris):
    L = lib()
    tris = np.ascontiguousarray(tris, np.int32)
    nb = L.orc_bending_quads(I(tris), len(tris), None)
    q = np.zeros((max(nb, 1), 4), np.int32)
    L.orc_bending_quads(I(tris), len(tris), I(q))
    return q[:nb]


def build_prims(pos, pairs, quads=None):
    """primList/indices/rest in the library's order: distance prims then bending."""
    L = lib()
    pos = np.ascontiguousarray(pos, np.float32)
    pairs = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    ne = len(pairs)
    rest_d = np.zeros(ne, np.float32)
    if ne:
        L.orc_rest_lengths(F(pos), I(pairs), ne, F(rest_d))
    quads = np.zeros((0, 4), np.int32) if quads is None else np.ascontiguousarray(quads, np.int32)
    nb = len(quads)
    rest_b = np.zeros(nb, np.float32)
    for i in range(nb):
        q = quads[i]
        rest_b[i] = L.orc_dihedral(F(pos[q[0]]), F(pos[q[1]]), F(pos[q[2]]), F(pos[q[3]]))
    prim_list = np.zeros((ne + nb, 2), np.int32)
    prim_list[:ne, 0] = 2 * np.arange(ne)
    prim_list[:ne, 1] = 2
    prim_list[ne:, 0] = 2 * ne + 4 * np.arange(nb)
    prim_list[ne:, 1] = 4
    indices = np.concatenate([pairs.reshape(-1), quads.reshape(-1)]).astype(np.int32)
    rest = np.concatenate([rest_d, rest_b]).astype(np.float32)
    return prim_list, indices, rest


def adjacency(prim_list, indices, nv):
    L = lib()
    prim_list = np.ascontiguousarray(prim_list, np.int32)
    indices = np.ascontiguousarray(indices, np.int32)
    npr = len(prim_list)
    tot = L.orc_prim_adjacency(I(prim_list), npr, I(indices), nv, None, None)
    sn = np.zeros((npr, 2), np.int32)
    adj = np.zeros(max(tot, 1), np.int32)
    L.orc_prim_adjacency(I(prim_list), npr, I(indices), nv, I(sn), I(adj))
    return sn, adj[:tot]


def color(prim_list, sn, adj, max_rounds=20000):
    L = lib()
    npr = len(prim_list)
    col = np.full(npr, -1, np.int32)
    prd = np.full(npr, -1, np.int32)
    adj = np.ascontiguousarray(adj, np.int32)
    if len(adj) == 0:
        adj = np.zeros(1, np.int32)
    rounds = L.orc_color(I(np.ascontiguousarray(prim_list, np.int32)), npr, I(np.ascontiguousarray(sn, np.int32)),
                         I(adj), max_rounds, I(col), I(prd))
    return col, rounds


def color_sort(col):
    L = lib()
    col = np.ascontiguousarray(col, np.int32)
    srt = np.zeros(len(col), np.int32)
    ws = np.zeros((int(col.max()) + 1 if len(col) else 1, 2), np.int32)
    C_ = L.orc_color_sort(I(col), len(col), I(srt), I(ws))
    return srt, ws[:C_]


class OracleCloth:
    """A cloth stepped by the oracle, built in the same order as the library."""

    def __init__(self, pos, tris, pairs, pins=(), bending=False, stiffness=(1.0, 1.0, 1.0), iterations=20,
                 quads=None, colors=None):
        """colors: per-prim colours computed elsewhere (the oracle's own
        colouring otherwise); the colour order is derived from them as usual."""
        self.L = lib()
        self.x = np.ascontiguousarray(pos, np.float32).copy()
        self.v = np.zeros_like(self.x)
        self.p = self.x.copy()
        self.nv = len(self.x)
        self.w = np.ones(self.nv, np.float32)
        self.w[list(pins)] = 0.0
        self.tris = np.ascontiguousarray(tris, np.int32)
        if bending and quads is None:
            quads = bending_quads(self.tris)
        self.prim_list, self.indices, self.rest = build_prims(self.x, pairs, quads if bending else None)
        if colors is None:
            self.sn, self.adj = adjacency(self.prim_list, self.indices, self.nv)
            self.color, self.rounds = color(self.prim_list, self.sn, self.adj)
        else:
            self.color, self.rounds = np.ascontiguousarray(colors, np.int32), -1
        self.sorted, self.worksets = color_sort(self.color)
        self.stiffness = stiffness
        self.set_iterations(iterations)
        self._mk_struct()

    def set_iterations(self, n):
        self.kmax = max(1, n)
        self.kd = np.zeros(self.kmax, np.float32)
        self.kb = np.zeros(self.kmax, np.float32)
        self.L.orc_stiffness_table_ref(self.stiffness[0], self.kmax, F(self.kd))
        self.L.orc_stiffness_table_const(self.stiffness[1], self.kmax, self.kmax, F(self.kb))
        if hasattr(self, "c"):
            self._mk_struct()

    def _mk_struct(self):
        c = OrcCloth()
        c.nv = self.nv
        c.x, c.v, c.p, c.w = F(self.x), F(self.v), F(self.p), F(self.w)
        c.np = len(self.prim_list)
        c.prim_list, c.indices, c.rest = I(self.prim_list), I(self.indices), F(self.rest)
        c.sorted, c.ncolors, c.worksets = I(self.sorted), len(self.worksets), I(self.worksets)
        c.k_dist, c.k_bend, c.k_len = F(self.kd), F(self.kb), self.kmax
        self.c = c

    @staticmethod
    def params(dt=1.0 / 600.0, gravity=(0.0, -9.8, 0.0), damping=1.0, wind=(0.0, 0.0, 0.0), wind_drag=0.0,
               sphere=None, box=None) -> OrcParams:
        p = OrcParams()
        p.dt = dt
        p.gravity[:] = gravity
        p.damp_factor = powf(damping, dt)
        p.wind[:] = wind
        p.wind_drag = wind_drag
        if sphere is not None:
            p.sphere_on = 1
            p.sphere_c[:] = sphere[0]
            p.sphere_r = sphere[1]
        if box is not None:
            p.box_on = 1
            p.box_lo[:] = box[0]
            p.box_hi[:] = box[1]
        return p

    def advect(self, prm: OrcParams):
        self.L.orc_advect(C.byref(prm), self.nv, F(self.w), F(self.x), F(self.v), F(self.p))

    def substep(self, prm: OrcParams, cap=20, check_every=1, tol=-1.0):
        """Advect + sweeps (fixed or to tolerance) + integrate; returns (sweeps, residual)."""
        self.advect(prm)
        r = C.c_double()
        n = self.L.orc_substep(C.byref(self.c), C.byref(prm), cap, check_every, tol, C.byref(r))
        return n, r.value

    def substep_omp(self, prm: OrcParams, cap=20, threads=1):
        """Advect + cap sweeps colour-parallel on `threads` host threads +
        integrate (equal to substep with the same cap, bit for bit)."""
        self.advect(prm)
        return self.L.orc_substep_omp(C.byref(self.c), C.byref(prm), cap, threads)

    def sweep(self, index):
        self.L.orc_sweep(C.byref(self.c), index)

    def substep_jacobi(self, prm: OrcParams, cap=20, omega=1.0, check_every=1, tol=-1.0):
        """Advect + Jacobi sweeps (fixed or to tolerance) + integrate; returns (sweeps, residual)."""
        self.advect(prm)
        r = C.c_double()
        n = self.L.orc_substep_jacobi(C.byref(self.c), C.byref(prm), cap, omega, check_every, tol, C.byref(r))
        return n, r.value

    def jacobi_sweep(self, index, omega=1.0, p=None):
        """One Jacobi sweep of p (default: the cloth's p); returns the new positions."""
        src = np.ascontiguousarray(self.p if p is None else p, np.float32).reshape(-1, 3)
        out = np.empty_like(src)
        self.L.orc_jacobi_sweep(C.byref(self.c), index, omega, F(src), F(out))
        return out

    def start64(self):
        """Starts an FP64 twin from the current FP32 state (x, v)."""
        self.x64 = self.x.astype(np.float64)
        self.v64 = self.v.astype(np.float64)
        self.p64 = self.x64.copy()

    def substeps64(self, prm: OrcParams, sweeps=20, n=1):
        """n plain substeps of the FP64 twin (same topology, order, rest values, stiffness)."""
        dp = C.POINTER(C.c_double)
        self.L.orc64_substeps.argtypes = [C.POINTER(OrcCloth), C.POINTER(OrcParams), dp, dp, dp, C.c_int, C.c_int]
        self.L.orc64_substeps.restype = None
        self.L.orc64_substeps(C.byref(self.c), C.byref(prm), self.x64.ctypes.data_as(dp),
                              self.v64.ctypes.data_as(dp), self.p64.ctypes.data_as(dp), sweeps, n)

    def residual(self, p=None):
        m, s, n = C.c_double(), C.c_double(), C.c_int()
        arr = self.p if p is None else np.ascontiguousarray(p, np.float32)
        self.L.orc_residual(C.byref(self.c), F(arr), C.byref(m), C.byref(s), C.byref(n))
        return m.value, (s.value / n.value) ** 0.5 if n.value else 0.0


# ---------------------------------------------------------------------------
# Broadphase / CCD
# ---------------------------------------------------------------------------
class OrcContact(C.Structure):
    _fields_ = [("v", C.c_int), ("f", C.c_int), ("t", C.c_float), ("n", C.c_float * 3), ("w", C.c_double * 4)]


def _sig_collide(L):
    L.orc_ccd_detect.restype = C.c_longlong
    L.orc_ccd_detect.argtypes = [C.POINTER(OrcGrid), _fp, _fp, _ip, C.c_int, C.c_int, C.c_float, C.c_longlong,
                                 C.POINTER(OrcContact)]
    L.orc_swept_pairs.restype = C.c_longlong
    L.orc_swept_pairs.argtypes = [_fp, _fp, _ip, C.c_int, C.c_int, C.c_float, C.c_int, C.c_longlong, _i64p]
    L.orc_ccd_detect_swept.restype = C.c_longlong
    L.orc_ccd_detect_swept.argtypes = [_fp, _fp, _ip, C.c_int, C.c_int, C.c_float, C.c_int, C.c_longlong,
                                       C.POINTER(OrcContact)]
    L.orc_vf_cross.restype = C.c_int
    L.orc_vf_cross.argtypes = [_fp] * 8 + [C.c_float]
    L.orc_crossings.restype = C.c_longlong
    L.orc_crossings.argtypes = [_fp, _fp, _ip, C.c_int, C.c_int, C.c_float, _fp, C.c_longlong, _i64p]
    L.orc_ccd_resolve.restype = None
    L.orc_ccd_resolve.argtypes = [C.POINTER(OrcContact), C.c_longlong, _ip, _fp, _fp, _fp, C.c_float, C.c_int,
                                  C.c_float]
    L.orc_edge_table.restype = C.c_int
    L.orc_edge_table.argtypes = [_ip, C.c_int, _ip, _ip]
    L.orc_ee_ccd.restype = C.c_int
    L.orc_ee_ccd.argtypes = [_fp] * 8 + [C.c_float, _fp, _fp, _dp]
    L.orc_ee_fix.restype = C.c_int
    L.orc_ee_fix.argtypes = [_fp] * 8 + [C.c_float, _fp]
    L.orc_ee_candidates.restype = C.c_longlong
    L.orc_ee_candidates.argtypes = [C.POINTER(OrcGrid), _fp, _ip, C.c_int, _ip, _ip, C.c_longlong, _ip]
    L.orc_ccd_detect_ee.restype = C.c_longlong
    L.orc_ccd_detect_ee.argtypes = [C.POINTER(OrcGrid), _fp, _fp, _ip, C.c_int, C.c_float, C.c_longlong, _ip, _fp,
                                    _fp, _dp]
    L.orc_ccd_resolve_ee.restype = None
    L.orc_ccd_resolve_ee.argtypes = [C.POINTER(OrcContact), C.c_longlong, _ip, C.c_longlong, _ip, _fp, _fp, _fp,
                                     C.c_int, C.c_float, C.c_int, C.c_float]


def grid(mode=0, centre=(0.0, 0.0, 0.0), size=(1, 1, 1), cell=(1.0, 1.0, 1.0), table_bits=16) -> OrcGrid:
    L = lib()
    g = OrcGrid()
    c = np.asarray(centre, np.float32)
    s = np.asarray(size, np.int32)
    cl = np.asarray(cell, np.float32)
    L.orc_grid_init(C.byref(g), F(c), I(s), F(cl))
    g.mode = mode
    g.table_bits = table_bits
    return g


def tri_keys(g, pos, tris):
    L = lib()
    pos = np.ascontiguousarray(pos, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    keys = np.zeros(len(tris), np.int32)
    ids = np.zeros(len(tris), np.int32)
    L.orc_tri_keys(C.byref(g), F(pos), I(tris), len(tris), I(keys), I(ids))
    return keys, ids


def sorted_keys(keys, ids):
    L = lib()
    k, i = keys.copy(), ids.copy()
    L.orc_stable_sort_pairs(I(k), I(i), len(k))
    return k, i


def cell_ranges(g, sk):
    L = lib()
    nc = L.orc_num_cells(C.byref(g))
    cs, ce = np.zeros(nc, np.int32), np.zeros(nc, np.int32)
    L.orc_cell_ranges(I(np.ascontiguousarray(sk, np.int32)), len(sk), nc, I(cs), I(ce))
    return cs, ce


def candidate_pairs(g, pos, tris, nv=None):
    L = lib()
    pos = np.ascontiguousarray(pos, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    nv = len(pos) if nv is None else nv
    n = L.orc_candidate_pairs(C.byref(g), F(pos), I(tris), len(tris), nv, 0, None)
    out = np.zeros(max(n, 1), np.int64)
    L.orc_candidate_pairs(C.byref(g), F(pos), I(tris), len(tris), nv, n, out.ctypes.data_as(_i64p))
    return out[:n]


def ccd_detect(g, x, p, tris, thickness):
    L = lib()
    _sig_collide(L)
    x = np.ascontiguousarray(x, np.float32)
    p = np.ascontiguousarray(p, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    n = L.orc_ccd_detect(C.byref(g), F(x), F(p), I(tris), len(tris), len(x), thickness, 0, None)
    arr = (OrcContact * max(n, 1))()
    L.orc_ccd_detect(C.byref(g), F(x), F(p), I(tris), len(tris), len(x), thickness, n, arr)
    return arr, n


def swept_pairs(x, p, tris, thickness, all_pairs=False):
    """Swept-box candidate keys (v << 32 | f), ascending; all_pairs drops the box test."""
    L = lib()
    _sig_collide(L)
    x = np.ascontiguousarray(x, np.float32)
    p = np.ascontiguousarray(p, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    n = L.orc_swept_pairs(F(x), F(p), I(tris), len(tris), len(x), thickness, int(all_pairs), 0, None)
    out = np.zeros(max(n, 1), np.int64)
    L.orc_swept_pairs(F(x), F(p), I(tris), len(tris), len(x), thickness, int(all_pairs), n, out.ctypes.data_as(_i64p))
    return out[:n]


def crossings(x, p, tris, thickness, w=None):
    """(v, f) pairs whose vertex crosses the face between x and p
    (NarrowVFCCDTest, the reference's post-resolve check), ascending; with w,
    pairs of four static (w == 0) vertices are skipped."""
    L = lib()
    _sig_collide(L)
    x = np.ascontiguousarray(x, np.float32)
    p = np.ascontiguousarray(p, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    wp = None if w is None else F(np.ascontiguousarray(w, np.float32))
    n = L.orc_crossings(F(x), F(p), I(tris), len(tris), len(x), thickness, wp, 0, None)
    out = np.zeros(max(n, 1), np.int64)
    L.orc_crossings(F(x), F(p), I(tris), len(tris), len(x), thickness, wp, n, out.ctypes.data_as(_i64p))
    return np.stack([out[:n] >> 32, out[:n] & 0xffffffff], 1)


def ccd_detect_swept(x, p, tris, thickness, all_pairs=False):
    """Narrowphase over the swept-box set (or over every pair)."""
    L = lib()
    _sig_collide(L)
    x = np.ascontiguousarray(x, np.float32)
    p = np.ascontiguousarray(p, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    n = L.orc_ccd_detect_swept(F(x), F(p), I(tris), len(tris), len(x), thickness, int(all_pairs), 0, None)
    arr = (OrcContact * max(n, 1))()
    L.orc_ccd_detect_swept(F(x), F(p), I(tris), len(tris), len(x), thickness, int(all_pairs), n, arr)
    return arr, n


def edge_table(tris):
    """Edges in first-sighting order, (lower, higher); per-triangle edge ids."""
    L = lib()
    _sig_collide(L)
    tris = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
    tedge = np.zeros(3 * max(len(tris), 1), np.int32)
    edges = np.zeros(6 * max(len(tris), 1), np.int32)
    ne = L.orc_edge_table(I(tris), len(tris), I(tedge), I(edges))
    return tedge[:3 * len(tris)], edges[:2 * ne].reshape(-1, 2)


def ee_candidates(g, p, tris):
    """Edge-edge candidate pairs (edge ids, ascending) and the edge table."""
    L = lib()
    _sig_collide(L)
    p = np.ascontiguousarray(p, np.float32)
    tris = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
    tedge, edges = edge_table(tris)
    ed = np.ascontiguousarray(edges.reshape(-1), np.int32)
    n = L.orc_ee_candidates(C.byref(g), F(p), I(tris), len(tris), I(tedge), I(ed), 0, None)
    out = np.zeros(2 * max(n, 1), np.int32)
    L.orc_ee_candidates(C.byref(g), F(p), I(tris), len(tris), I(tedge), I(ed), n, I(out))
    return out[:2 * n].reshape(-1, 2), edges


def ee_ccd(x4, p4, thickness):
    """One edge pair: x4, p4 = (a0, a1, b0, b1) start and end positions.
    Returns (hit, t, n, w)."""
    L = lib()
    _sig_collide(L)
    xs = [np.ascontiguousarray(v, np.float32) for v in x4]
    ps = [np.ascontiguousarray(v, np.float32) for v in p4]
    t = np.zeros(1, np.float32)
    n = np.zeros(3, np.float32)
    w = np.zeros(4, np.float64)
    hit = L.orc_ee_ccd(*[F(v) for v in xs], *[F(v) for v in ps], thickness, F(t), F(n), w.ctypes.data_as(_dp))
    return bool(hit), float(t[0]), n, w


def ee_fix(X4, q4, thickness):
    """Displacements (4 x 3) of the edge-edge response, or None."""
    L = lib()
    _sig_collide(L)
    xs = [np.ascontiguousarray(v, np.float32) for v in X4]
    qs = [np.ascontiguousarray(v, np.float32) for v in q4]
    d = np.zeros(12, np.float32)
    ok = L.orc_ee_fix(*[F(v) for v in xs], *[F(v) for v in qs], thickness, F(d))
    return d.reshape(4, 3) if ok else None


def ccd_detect_ee(g, x, p, tris, thickness):
    """Edge-edge contacts: quads (a0, a1, b0, b1), t, n, w, in candidate order."""
    L = lib()
    _sig_collide(L)
    x = np.ascontiguousarray(x, np.float32)
    p = np.ascontiguousarray(p, np.float32)
    tris = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
    n = L.orc_ccd_detect_ee(C.byref(g), F(x), F(p), I(tris), len(tris), thickness, 0, None, None, None, None)
    m = max(n, 1)
    q = np.zeros((m, 4), np.int32)
    t = np.zeros(m, np.float32)
    nn = np.zeros((m, 3), np.float32)
    w = np.zeros((m, 4), np.float64)
    L.orc_ccd_detect_ee(C.byref(g), F(x), F(p), I(tris), len(tris), thickness, n, I(q), F(t), F(nn),
                        w.ctypes.data_as(_dp))
    return q[:n], t[:n], nn[:n], w[:n]


def ccd_resolve_ee(arr, n, quads, tris, x, p, w, thickness, passes, min_n2=1e-6):
    """Vertex-face contacts (arr, n) and edge-edge contacts (quads) resolved
    together, `passes` Jacobi passes."""
    L = lib()
    _sig_collide(L)
    p = np.ascontiguousarray(p, np.float32).copy()
    q = np.ascontiguousarray(quads, np.int32).reshape(-1, 4)
    if arr is None:
        arr, n = (OrcContact * 1)(), 0
    L.orc_ccd_resolve_ee(arr, n, I(q), len(q), I(np.ascontiguousarray(tris, np.int32)),
                         F(np.ascontiguousarray(x, np.float32)), F(p), F(np.ascontiguousarray(w, np.float32)),
                         len(p), thickness, passes, min_n2)
    return p


def contacts_arrays(arr, n):
    vf = np.array([[arr[i].v, arr[i].f] for i in range(n)], np.int32).reshape(-1, 2)
    t = np.array([arr[i].t for i in range(n)], np.float32)
    nn = np.array([list(arr[i].n) for i in range(n)], np.float32).reshape(-1, 3)
    w = np.array([list(arr[i].w) for i in range(n)], np.float64).reshape(-1, 4)
    return vf, t, nn, w


def ccd_resolve(arr, n, tris, x, p, w, thickness, passes, min_n2=1e-6):
    """min_n2: degenerate-face bound of VFResolveTest (1e-6 in the reference)."""
    L = lib()
    _sig_collide(L)
    p = np.ascontiguousarray(p, np.float32).copy()
    L.orc_ccd_resolve(arr, n, I(np.ascontiguousarray(tris, np.int32)), F(np.ascontiguousarray(x, np.float32)),
                      F(p), F(np.ascontiguousarray(w, np.float32)), thickness, passes, min_n2)
    return p
