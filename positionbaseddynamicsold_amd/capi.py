"""ctypes binding of libpbd.so (the C ABI declared in include/pbd.h).

The library is loaded from this package directory; a missing or unloadable
library raises immediately -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libpbd.so"

PBD_OK = 0
PBD_ERR_ARG = -1
PBD_ERR_STATE = -2
PBD_ERR_HIP = -3
PBD_ERR_LIMIT = -4
PBD_ERR_IO = -5
PBD_ERR_COLOR = -6

DISTANCE = 0x001
BENDING = 0x010
ANCHOR = 0x100

STATE_X, STATE_V, STATE_P = 0, 1, 2
PATH_AUTO, PATH_GLOBAL, PATH_JACOBI = 0, 1, 2


class GridDesc(C.Structure):
    _fields_ = [("res_x", C.c_int), ("res_y", C.c_int), ("size_x", C.c_float), ("size_y", C.c_float),
                ("origin", C.c_float * 3), ("jitter", C.c_float), ("seed", C.c_uint64)]


class ClothParams(C.Structure):
    _fields_ = [("damping", C.c_float), ("gravity", C.c_float * 3), ("wind", C.c_float * 3),
                ("wind_drag", C.c_float), ("sphere_on", C.c_int), ("sphere_c", C.c_float * 3),
                ("sphere_r", C.c_float), ("box_on", C.c_int), ("box_lo", C.c_float * 3),
                ("box_hi", C.c_float * 3)]


class SolverParams(C.Structure):
    _fields_ = [("dt", C.c_float), ("iterations", C.c_int), ("tol", C.c_double), ("check_every", C.c_int),
                ("path", C.c_int), ("tile_sweeps", C.c_int), ("tile_max_verts", C.c_int),
                ("block_threads", C.c_int), ("graph", C.c_int), ("warm_iterations", C.c_int),
                ("warm_rule", C.c_int), ("warm_gate", C.c_float), ("jacobi_omega", C.c_float),
                ("split_whole", C.c_int), ("per_chunk_launches", C.c_int)]


class HashDesc(C.Structure):
    _fields_ = [("mode", C.c_int), ("centre", C.c_float * 3), ("size", C.c_int * 3), ("cell", C.c_float * 3),
                ("table_bits", C.c_int)]


def hash_desc(mode=0, centre=(0.0, 0.0, 0.0), size=(1, 1, 1), cell=(1.0, 1.0, 1.0), table_bits=16) -> HashDesc:
    return HashDesc(mode, (C.c_float * 3)(*centre), (C.c_int * 3)(*size), (C.c_float * 3)(*cell), table_bits)


class CollisionParams(C.Structure):
    _fields_ = [("enabled", C.c_int), ("grid", HashDesc), ("cell_scale", C.c_float), ("thickness", C.c_float),
                ("resolve_passes", C.c_int), ("detect_every", C.c_int), ("min_normal2", C.c_float),
                ("broadphase", C.c_int), ("edge_edge", C.c_int)]


class Profile(C.Structure):
    _fields_ = [("tile_ms", C.c_double), ("tile_launches", C.c_longlong), ("tile_alg_bytes", C.c_double),
                ("step_ms", C.c_double), ("substeps", C.c_longlong), ("vertex_sweeps", C.c_longlong),
                ("sweeps_max", C.c_longlong), ("graph_replays", C.c_longlong), ("ccd_detections", C.c_longlong),
                ("ccd_triangles", C.c_longlong), ("ccd_cell_entries", C.c_longlong), ("ccd_raw_pairs", C.c_longlong),
                ("ccd_pairs", C.c_longlong), ("ccd_contacts", C.c_longlong), ("ccd_resolved", C.c_longlong),
                ("ccd_ee_raw_pairs", C.c_longlong), ("ccd_ee_pairs", C.c_longlong), ("ccd_ee_contacts", C.c_longlong),
                ("tile_bytes", C.c_double), ("queue_wait_timeouts", C.c_longlong),
                ("queue_deciding_launches", C.c_longlong)]


class ClothInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nv", "nt", "ne", "nb", "nprims", "ncolors", "color_rounds", "adj_len",
                                       "ntiles", "whole", "max_loaded")]


_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)

# name -> argtypes (all return int except pbd_last_error)
SIGNATURES = {
    "pbd_create": [_I, C.POINTER(_P)],
    "pbd_destroy": [_P],
    "pbd_synchronize": [_P],
    "pbd_cloth_create_grid": [_P, C.POINTER(GridDesc), _ip],
    "pbd_cloth_create_mesh": [_P, _fp, _I, _ip, _I, _ip],
    "pbd_cloth_load_txt": [_P, C.c_char_p, C.c_char_p, _ip],
    "pbd_cloth_load_cache": [_P, C.c_char_p, C.c_char_p, _ip],
    "pbd_ccd_detect_ee": [_P, C.POINTER(HashDesc), _fp, _fp, _I, _ip, _I, C.c_float, C.POINTER(C.c_longlong),
                          C.POINTER(C.c_longlong)],
    "pbd_ccd_get_ee_pairs": [_P, C.c_longlong, _ip, C.POINTER(C.c_longlong)],
    "pbd_ccd_get_ee_contacts": [_P, C.c_longlong, _ip, _fp, _fp, C.POINTER(C.c_double), C.POINTER(C.c_longlong)],
    "pbd_cloth_save_cache": [_P, _I, C.c_char_p, C.c_char_p, _I],
    "pbd_cloth_set_distance_pairs": [_P, _I, _ip, _I],
    "pbd_cloth_set_constraint_option": [_P, _I, C.c_uint, _fp],
    "pbd_cloth_set_pins": [_P, _I, _ip, _I],
    "pbd_cloth_set_params": [_P, _I, C.POINTER(ClothParams)],
    "pbd_cloth_get_params": [_P, _I, C.POINTER(ClothParams)],
    "pbd_init": [_P, _I],
    "pbd_cloth_get_info": [_P, _I, C.POINTER(ClothInfo)],
    "pbd_cloth_get_prims": [_P, _I, _ip, _ip, _fp],
    "pbd_cloth_get_adjacency": [_P, _I, _ip, _ip],
    "pbd_cloth_get_colors": [_P, _I, _ip, _ip, _ip],
    "pbd_set_solver_params": [_P, C.POINTER(SolverParams)],
    "pbd_advect": [_P, _F],
    "pbd_project": [_P, _I],
    "pbd_integrate": [_P, _F],
    "pbd_step": [_P, _I],
    "pbd_step_cloth": [_P, _I, _I],
    "pbd_autotune": [_P, _I, C.POINTER(SolverParams), _ip, _dp, _ip],
    "pbd_cloth_get_state": [_P, _I, _I, _fp],
    "pbd_cloth_set_state": [_P, _I, _I, _fp],
    "pbd_cloth_residual": [_P, _I, _dp, _dp],
    "pbd_cloth_last_sweeps": [_P, _I, _ip],
    "pbd_sh_build": [_P, C.POINTER(HashDesc), _fp, _I, _ip, _I],
    "pbd_sh_build_noalloc": [_I, _P, C.POINTER(HashDesc), _P, _I, _P, _I, _P, _P, _P, _P, _P],
    "pbd_sh_num_cells": [_P, _ip],
    "pbd_sh_get": [_P, _ip, _ip, _ip, _ip, _ip],
    "pbd_sh_pairs": [_P, C.c_longlong, C.POINTER(C.c_int64), C.POINTER(C.c_longlong)],
    "pbd_ccd_detect": [_P, C.POINTER(HashDesc), _fp, _fp, _I, _ip, _I, _F, C.POINTER(C.c_longlong)],
    "pbd_ccd_get_contacts": [_P, C.c_longlong, _ip, _fp, _fp, _dp],
    "pbd_ccd_detect_swept": [_P, C.POINTER(HashDesc), _fp, _fp, _I, _ip, _I, _F, C.POINTER(C.c_longlong),
                             C.POINTER(C.c_longlong)],
    "pbd_ccd_get_pairs": [_P, C.c_longlong, C.POINTER(C.c_int64), C.POINTER(C.c_longlong)],
    "pbd_ccd_resolve": [_P, _fp, _fp, _F, _I],
    "pbd_set_collision": [_P, C.POINTER(CollisionParams)],
    "pbd_last_contacts": [_P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)],
    "pbd_ws_num_params": [_ip],
    "pbd_ws_init": [_P, C.c_uint64],
    "pbd_ws_get_params": [_P, _fp],
    "pbd_ws_set_params": [_P, _fp],
    "pbd_ws_collect": [_P, _I, _I],
    "pbd_ws_num_samples": [_P, C.POINTER(C.c_longlong)],
    "pbd_ws_get_samples": [_P, C.c_longlong, C.c_longlong, _fp, _fp],
    "pbd_ws_set_samples": [_P, _fp, _fp, C.c_longlong],
    "pbd_ws_forward": [_P, _fp, _I, _fp],
    "pbd_ws_gradient": [_P, C.c_longlong, _I, _fp, _dp],
    "pbd_ws_adam_step": [_P, _fp, _F],
    "pbd_ws_train": [_P, _I, _I, _F, C.c_uint64, _dp],
    "pbd_ws_enable": [_P, _I],
    "pbd_ws_get_normalization": [_P, _fp],
    "pbd_ws_set_normalization": [_P, _fp],
    "pbd_cloth_set_tolerance": [_P, _I, C.c_double],
    "pbd_residual_all": [_P, _dp, _dp],
    "pbd_num_vertices": [_P, _ip],
    "pbd_state_save": [_P, _fp, _fp],
    "pbd_state_load": [_P, _fp, _fp],
    "pbd_snapshot": [_P, _I],
    "pbd_ws_last_used": [_P, _I, _ip],
    "pbd_ws_last_delta": [_P, _I, _fp],
    "pbd_profile_enable": [_P, _I],
    "pbd_profile_read": [_P, C.POINTER(Profile), _I],
    "pbd_bench_copy": [_P, C.c_longlong, _I, _dp],
}

_lib = None


def load() -> C.CDLL:
    """Loads libpbd.so once; raises if it is missing (build it first)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run positionbaseddynamicsold_amd.build.build() first")
        lib = C.CDLL(str(LIB_PATH))
        lib.pbd_last_error.restype = C.c_char_p
        lib.pbd_last_error.argtypes = []
        for name, args in SIGNATURES.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = C.c_int
        _lib = lib
    return _lib


class PbdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pbd error {code}: {msg}")
        self.code = code


def check(code: int) -> None:
    if code != PBD_OK:
        raise PbdError(code, load().pbd_last_error().decode(errors="replace"))


def fptr(a: np.ndarray):
    return a.ctypes.data_as(_fp)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def default_cloth_params(**kw) -> ClothParams:
    p = ClothParams()
    p.damping = 1.0
    p.gravity[:] = (0.0, -9.8, 0.0)
    for k, v in kw.items():
        if isinstance(getattr(p, k), C.Array):
            getattr(p, k)[:] = v
        else:
            setattr(p, k, v)
    return p


class Context:
    """One pbd_ctx: a batch of cloths on one device."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.h = C.c_void_p()
        check(self.lib.pbd_create(device, C.byref(self.h)))
        self.cloths: list[int] = []

    def close(self):
        if self.h:
            check(self.lib.pbd_destroy(self.h))
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- model ---------------------------------------------------------
    def add_grid(self, rx, ry, sx=1.0, sy=1.0, origin=(0.0, 0.0, 0.0), jitter=0.0, seed=0) -> int:
        d = GridDesc(rx, ry, sx, sy, (C.c_float * 3)(*origin), jitter, seed)
        cid = C.c_int()
        check(self.lib.pbd_cloth_create_grid(self.h, C.byref(d), C.byref(cid)))
        self.cloths.append(cid.value)
        return cid.value

    def add_mesh(self, pos: np.ndarray, tris: np.ndarray) -> int:
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        tris = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
        cid = C.c_int()
        check(self.lib.pbd_cloth_create_mesh(self.h, fptr(pos), len(pos), iptr(tris), len(tris), C.byref(cid)))
        self.cloths.append(cid.value)
        return cid.value

    def load_txt(self, topol: str, constr: str | None = None) -> int:
        cid = C.c_int()
        check(self.lib.pbd_cloth_load_txt(self.h, topol.encode(), (constr or "").encode(), C.byref(cid)))
        self.cloths.append(cid.value)
        return cid.value

    def load_cache(self, mesh: str, constr: str | None = None) -> int:
        """PBDObject::ContinueSimInit (PBD_cuda.cu:660-667): a cloth continued
        from the reference's mesh and constraint cache files."""
        cid = C.c_int()
        check(self.lib.pbd_cloth_load_cache(self.h, mesh.encode(), constr.encode() if constr else None,
                                            C.byref(cid)))
        self.cloths.append(cid.value)
        return cid.value

    def save_cache(self, cloth: int, mesh: str | None, constr: str | None = None, precision: int = 0):
        """PBDObject::SaveMeshTopol + SaveConstraint (PBD_cuda.cu:907-938);
        precision 0 writes the reference's 6 significant digits, 9 is exact."""
        check(self.lib.pbd_cloth_save_cache(self.h, cloth, mesh.encode() if mesh else None,
                                            constr.encode() if constr else None, precision))

    def set_distance_pairs(self, cloth: int, pairs: np.ndarray):
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        check(self.lib.pbd_cloth_set_distance_pairs(self.h, cloth, iptr(pairs), len(pairs)))

    def set_constraint_option(self, cloth: int, ct: int, stiffness=(1.0, 1.0, 1.0)):
        st = np.asarray(stiffness, dtype=np.float32)
        check(self.lib.pbd_cloth_set_constraint_option(self.h, cloth, ct, fptr(st)))

    def set_pins(self, cloth: int, pins):
        pins = np.ascontiguousarray(pins, dtype=np.int32)
        check(self.lib.pbd_cloth_set_pins(self.h, cloth, iptr(pins), len(pins)))

    def set_params(self, cloth: int, params: ClothParams | None = None, **kw):
        p = params if params is not None else default_cloth_params(**kw)
        check(self.lib.pbd_cloth_set_params(self.h, cloth, C.byref(p)))

    def get_params(self, cloth: int) -> ClothParams:
        p = ClothParams()
        check(self.lib.pbd_cloth_get_params(self.h, cloth, C.byref(p)))
        return p

    def update_params(self, cloth: int, **kw):
        """Changes the named fields, keeping the others (e.g. a cache's damping and gravity)."""
        p = self.get_params(cloth)
        for k, v in kw.items():
            f = getattr(p, k)
            if hasattr(f, "__len__"):
                f[:] = v
            else:
                setattr(p, k, v)
        self.set_params(cloth, p)

    def init(self, color_rounds: int = 20000):
        check(self.lib.pbd_init(self.h, color_rounds))

    def info(self, cloth: int) -> ClothInfo:
        i = ClothInfo()
        check(self.lib.pbd_cloth_get_info(self.h, cloth, C.byref(i)))
        return i

    def prims(self, cloth: int):
        inf = self.info(cloth)
        pl = np.zeros((inf.nprims, 2), np.int32)
        idx = np.zeros(2 * inf.ne + 4 * inf.nb, np.int32)
        rest = np.zeros(inf.nprims, np.float32)
        check(self.lib.pbd_cloth_get_prims(self.h, cloth, iptr(pl), iptr(idx), fptr(rest)))
        return pl, idx, rest

    def adjacency(self, cloth: int):
        inf = self.info(cloth)
        sn = np.zeros((inf.nprims, 2), np.int32)
        adj = np.zeros(max(1, inf.adj_len), np.int32)
        check(self.lib.pbd_cloth_get_adjacency(self.h, cloth, iptr(sn), iptr(adj)))
        return sn, adj[: inf.adj_len]

    def colors(self, cloth: int):
        inf = self.info(cloth)
        col = np.zeros(inf.nprims, np.int32)
        srt = np.zeros(inf.nprims, np.int32)
        ws = np.zeros((max(1, inf.ncolors), 2), np.int32)
        check(self.lib.pbd_cloth_get_colors(self.h, cloth, iptr(col), iptr(srt), iptr(ws)))
        return col, srt, ws[: inf.ncolors]

    # ---- solver --------------------------------------------------------
    def set_solver(self, dt=1.0 / 600.0, iterations=20, tol=-1.0, check_every=1, path=PATH_AUTO,
                   tile_sweeps=0, tile_max_verts=0, block_threads=0, graph=0, warm_iterations=0, warm_rule=0,
                   warm_gate=0.0, jacobi_omega=0.0, split_whole=0, per_chunk_launches=0):
        p = SolverParams(dt, iterations, tol, check_every, path, tile_sweeps, tile_max_verts, block_threads, graph,
                         warm_iterations, warm_rule, warm_gate, jacobi_omega, split_whole, per_chunk_launches)
        check(self.lib.pbd_set_solver_params(self.h, C.byref(p)))

    def set_solver_params(self, p: SolverParams):
        check(self.lib.pbd_set_solver_params(self.h, C.byref(p)))

    def autotune(self, trial_substeps: int = 3):
        """Picks block size, sweeps per launch, tile size (and hash table size
        with collisions) by timing; returns (params, table_bits, ms, trials)."""
        p = SolverParams()
        bits, ms, n = C.c_int(), C.c_double(), C.c_int()
        check(self.lib.pbd_autotune(self.h, trial_substeps, C.byref(p), C.byref(bits), C.byref(ms), C.byref(n)))
        return p, bits.value, ms.value, n.value

    def advect(self, dt: float):
        check(self.lib.pbd_advect(self.h, dt))

    def project(self, iterations: int):
        check(self.lib.pbd_project(self.h, iterations))

    def integrate(self, dt: float):
        check(self.lib.pbd_integrate(self.h, dt))

    def step(self, substeps: int = 1):
        check(self.lib.pbd_step(self.h, substeps))

    def step_cloth(self, cloth: int, substeps: int = 1):
        """Substeps of one cloth alone (the unbatched solver)."""
        check(self.lib.pbd_step_cloth(self.h, cloth, substeps))

    def sync(self):
        check(self.lib.pbd_synchronize(self.h))

    def get(self, cloth: int, which: int = STATE_X) -> np.ndarray:
        n = self.info(cloth).nv
        out = np.zeros((n, 3), np.float32)
        check(self.lib.pbd_cloth_get_state(self.h, cloth, which, fptr(out)))
        return out

    def set(self, cloth: int, which: int, xyz: np.ndarray):
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        check(self.lib.pbd_cloth_set_state(self.h, cloth, which, fptr(xyz)))

    def residual(self, cloth: int):
        m, r = C.c_double(), C.c_double()
        check(self.lib.pbd_cloth_residual(self.h, cloth, C.byref(m), C.byref(r)))
        return m.value, r.value

    # ---- broadphase / CCD ---------------------------------------------
    def sh_build(self, desc: HashDesc, pos: np.ndarray, tris: np.ndarray):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        tris = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
        check(self.lib.pbd_sh_build(self.h, C.byref(desc), fptr(pos), len(pos), iptr(tris), len(tris)))
        self._sh_nt = len(tris)

    def sh_get(self):
        nc = C.c_int()
        check(self.lib.pbd_sh_num_cells(self.h, C.byref(nc)))
        nt = self._sh_nt
        keys, ids, raw = (np.zeros(max(nt, 1), np.int32) for _ in range(3))
        cs, ce = np.zeros(nc.value, np.int32), np.zeros(nc.value, np.int32)
        check(self.lib.pbd_sh_get(self.h, iptr(keys), iptr(ids), iptr(raw), iptr(cs), iptr(ce)))
        return keys[:nt], ids[:nt], raw[:nt], cs, ce

    def sh_pairs(self) -> np.ndarray:
        n = C.c_longlong()
        check(self.lib.pbd_sh_pairs(self.h, 0, None, C.byref(n)))
        out = np.zeros(max(n.value, 1), np.int64)
        check(self.lib.pbd_sh_pairs(self.h, n.value, out.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(n)))
        return out[: n.value]

    def ccd_detect(self, desc: HashDesc, x, p, tris, thickness: float, swept: bool = False):
        """Contacts (vf, t, n, w) of one detection; swept=True uses the
        swept-box candidate set (its pairs: ccd_pairs())."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
        p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
        tris = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
        n = C.c_longlong()
        if swept:
            npairs = C.c_longlong()
            check(self.lib.pbd_ccd_detect_swept(self.h, C.byref(desc), fptr(x), fptr(p), len(x), iptr(tris),
                                                len(tris), thickness, C.byref(npairs), C.byref(n)))
        else:
            check(self.lib.pbd_ccd_detect(self.h, C.byref(desc), fptr(x), fptr(p), len(x), iptr(tris), len(tris),
                                          thickness, C.byref(n)))
        m = n.value
        vf = np.zeros((max(m, 1), 2), np.int32)
        t = np.zeros(max(m, 1), np.float32)
        nn = np.zeros((max(m, 1), 3), np.float32)
        w = np.zeros((max(m, 1), 4), np.float64)
        check(self.lib.pbd_ccd_get_contacts(self.h, m, iptr(vf), fptr(t), fptr(nn), w.ctypes.data_as(_dp)))
        return vf[:m], t[:m], nn[:m], w[:m]

    def ccd_detect_ee(self, desc: HashDesc, x, p, tris, thickness: float):
        """Edge-edge contacts (quads (a0, a1, b0, b1), t, n, w) of one detection,
        added to those of the preceding ccd_detect on the same scene."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
        p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
        tris = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
        n, npairs = C.c_longlong(), C.c_longlong()
        check(self.lib.pbd_ccd_detect_ee(self.h, C.byref(desc), fptr(x), fptr(p), len(x), iptr(tris), len(tris),
                                         thickness, C.byref(npairs), C.byref(n)))
        m = n.value
        q = np.zeros((max(m, 1), 4), np.int32)
        t = np.zeros(max(m, 1), np.float32)
        nn = np.zeros((max(m, 1), 3), np.float32)
        w = np.zeros((max(m, 1), 4), np.float64)
        cnt = C.c_longlong()
        check(self.lib.pbd_ccd_get_ee_contacts(self.h, m, iptr(q), fptr(t), fptr(nn), w.ctypes.data_as(_dp),
                                               C.byref(cnt)))
        return q[:m], t[:m], nn[:m], w[:m]

    def ccd_ee_pairs(self) -> np.ndarray:
        """Edge-edge candidates of the last ccd_detect_ee, (a0, a1, b0, b1) rows."""
        cnt = C.c_longlong()
        check(self.lib.pbd_ccd_get_ee_pairs(self.h, 0, None, C.byref(cnt)))
        q = np.zeros((max(cnt.value, 1), 4), np.int32)
        check(self.lib.pbd_ccd_get_ee_pairs(self.h, cnt.value, iptr(q), C.byref(cnt)))
        return q[:cnt.value]

    def ccd_pairs(self) -> np.ndarray:
        """Candidate pairs of the last detection as sorted keys v << 32 | f."""
        n = C.c_longlong()
        check(self.lib.pbd_ccd_get_pairs(self.h, 0, None, C.byref(n)))
        out = np.zeros(max(n.value, 1), np.int64)
        check(self.lib.pbd_ccd_get_pairs(self.h, n.value, out.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(n)))
        return out[: n.value]

    def ccd_resolve(self, p, w, thickness: float, passes: int) -> np.ndarray:
        p = np.ascontiguousarray(p, np.float32).reshape(-1, 3).copy()
        w = np.ascontiguousarray(w, np.float32)
        check(self.lib.pbd_ccd_resolve(self.h, fptr(p), fptr(w), thickness, passes))
        return p

    def set_collision(self, enabled=1, grid: HashDesc | None = None, cell_scale=1.5, thickness=0.005,
                      resolve_passes=2, detect_every=10, min_normal2=0.0, broadphase=0, edge_edge=0):
        """min_normal2 = 0 keeps the reference's degenerate-face bound (1e-6 m^4);
        broadphase 0 = the reference's centroid cells, 1 = swept boxes;
        edge_edge 1 adds edge-edge contacts."""
        cp = CollisionParams(enabled, grid if grid is not None else hash_desc(1, table_bits=20), cell_scale,
                             thickness, resolve_passes, detect_every, min_normal2, broadphase, edge_edge)
        check(self.lib.pbd_set_collision(self.h, C.byref(cp)))

    def last_contacts(self):
        a, b = C.c_longlong(), C.c_longlong()
        check(self.lib.pbd_last_contacts(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    # ---- learned warm start --------------------------------------------
    @staticmethod
    def ws_num_params() -> int:
        n = C.c_int()
        check(load().pbd_ws_num_params(C.byref(n)))
        return n.value

    def ws_init(self, seed: int = 0):
        check(self.lib.pbd_ws_init(self.h, seed))

    def ws_params(self) -> np.ndarray:
        out = np.zeros(self.ws_num_params(), np.float32)
        check(self.lib.pbd_ws_get_params(self.h, fptr(out)))
        return out

    def ws_set_params(self, prm: np.ndarray):
        prm = np.ascontiguousarray(prm, np.float32)
        check(self.lib.pbd_ws_set_params(self.h, fptr(prm)))

    def ws_collect(self, sweeps: int, substeps: int):
        check(self.lib.pbd_ws_collect(self.h, sweeps, substeps))

    def ws_num_samples(self) -> int:
        n = C.c_longlong()
        check(self.lib.pbd_ws_num_samples(self.h, C.byref(n)))
        return n.value

    def ws_samples(self, first: int = 0, n: int | None = None):
        n = self.ws_num_samples() - first if n is None else n
        X = np.zeros((n, 22), np.float32)
        Y = np.zeros((n, 3), np.float32)
        check(self.lib.pbd_ws_get_samples(self.h, first, n, fptr(X), fptr(Y)))
        return X, Y

    def ws_set_samples(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        check(self.lib.pbd_ws_set_samples(self.h, fptr(X), fptr(Y), len(X)))

    def ws_forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32).reshape(-1, 22)
        Y = np.zeros((len(X), 3), np.float32)
        check(self.lib.pbd_ws_forward(self.h, fptr(X), len(X), fptr(Y)))
        return Y

    def ws_gradient(self, first: int, n: int):
        g = np.zeros(self.ws_num_params(), np.float32)
        loss = C.c_double()
        check(self.lib.pbd_ws_gradient(self.h, first, n, fptr(g), C.byref(loss)))
        return g, loss.value

    def ws_adam_step(self, grad, lr: float):
        grad = np.ascontiguousarray(grad, np.float32)
        check(self.lib.pbd_ws_adam_step(self.h, fptr(grad), lr))

    def ws_train(self, epochs: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0) -> float:
        loss = C.c_double()
        check(self.lib.pbd_ws_train(self.h, epochs, batch, lr, seed, C.byref(loss)))
        return loss.value

    def ws_norm(self) -> np.ndarray:
        out = np.zeros(47, np.float32)
        check(self.lib.pbd_ws_get_normalization(self.h, fptr(out)))
        return out

    def ws_set_norm(self, norm):
        norm = np.ascontiguousarray(norm, np.float32)
        check(self.lib.pbd_ws_set_normalization(self.h, fptr(norm)))

    def residual_all(self):
        n = len(self.cloths)
        m, r = np.zeros(n), np.zeros(n)
        check(self.lib.pbd_residual_all(self.h, m.ctypes.data_as(_dp), r.ctypes.data_as(_dp)))
        return m, r

    def num_vertices(self) -> int:
        n = C.c_int()
        check(self.lib.pbd_num_vertices(self.h, C.byref(n)))
        return n.value

    def state_save(self):
        n = self.num_vertices()
        x, v = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
        check(self.lib.pbd_state_save(self.h, fptr(x), fptr(v)))
        return x, v

    def state_load(self, x, v):
        x = np.ascontiguousarray(x, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        check(self.lib.pbd_state_load(self.h, fptr(x), fptr(v)))

    def snapshot(self, restore: bool = False):
        check(self.lib.pbd_snapshot(self.h, 1 if restore else 0))

    def set_tolerance(self, cloth: int, tol: float):
        check(self.lib.pbd_cloth_set_tolerance(self.h, cloth, tol))

    def ws_enable(self, on: bool = True):
        check(self.lib.pbd_ws_enable(self.h, 1 if on else 0))

    def ws_last_delta(self, cloth: int) -> np.ndarray:
        out = np.zeros((self.info(cloth).nv, 3), np.float32)
        check(self.lib.pbd_ws_last_delta(self.h, cloth, fptr(out)))
        return out

    def ws_last_used(self, cloth: int) -> int:
        u = C.c_int()
        check(self.lib.pbd_ws_last_used(self.h, cloth, C.byref(u)))
        return u.value

    def profile(self, on: bool = True):
        check(self.lib.pbd_profile_enable(self.h, 1 if on else 0))

    def profile_read(self, reset: bool = True) -> Profile:
        p = Profile()
        check(self.lib.pbd_profile_read(self.h, C.byref(p), 1 if reset else 0))
        return p

    def bench_copy(self, nbytes: int, reps: int = 20) -> float:
        ms = C.c_double()
        check(self.lib.pbd_bench_copy(self.h, nbytes, reps, C.byref(ms)))
        return ms.value

    def last_sweeps(self, cloth: int) -> int:
        s = C.c_int()
        check(self.lib.pbd_cloth_last_sweeps(self.h, cloth, C.byref(s)))
        return s.value
