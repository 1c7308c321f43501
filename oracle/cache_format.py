"""The reference's text cache format, restated in Python -- TEST INFRASTRUCTURE.

Only tests/ import this module; it never participates in the product path.

Writer: Buffer<T>::Save (Buffer.cuh:70-80) writes one line per buffer,
``name|type|item;item;...;`` followed by std::endl; IO::SaveData
(PBD_Basic.cuh:278-296) writes ``name|type|item`` (no ';').  A float item is
``ofstream << float`` with the default stream flags, which formats the value
widened to double as printf's %g at 6 significant digits; float3 and int2 items
are their components joined by ',' (PBD_Basic.cuh:68-78).  The mesh file is
PBDObject::SaveMeshTopol (PBD_cuda.cu:907-925): Indices, P, primList,
velBuffer, massBuffer, dampingRate, gravity.  The constraint file is
PBDObject::SaveConstraint -> ConstraintPBD::Save (PBD_cuda.cu:627-640,
927-938): Indices, P, primList, prdPBuffer, restPosBuffer, restLengthBuffer,
stiffnessBuffer, constraintType.

Reader: IO::ReadTopolFromCache / IO::ReadConstraintFromCache
(PBD_Basic.cuh:326-530): every line is split on '|' with UT::splitString
(Utility.h:8-28, which drops an empty last piece); lines without exactly three
pieces are skipped; the third piece is split on ';' and each item on ','.

Parity: the %g rule is pinned against the C++ stream itself (a helper built
with g++ in tests/test_cache_format.py); the reader's split rule by the cases
the reference's splitString admits (trailing separators, empty buffers, CRLF).
"""
from __future__ import annotations

import numpy as np

DISTANCE, BENDING, ANCHOR = 0x001, 0x010, 0x100
# typeid(T).name() as the reference's compiler (MSVC, PositionBasedDynamicsCUDA.vcxproj) spells it
T_INT, T_FLOAT, T_FLOAT3, T_INT2 = "int", "float", "struct float3", "struct int2"


def split(s: str, sep: str) -> list[str]:
    """UT::splitString (Utility.h:8-28)."""
    out, a = [], 0
    while True:
        b = s.find(sep, a)
        if b < 0:
            break
        out.append(s[a:b])
        a = b + len(sep)
    if a != len(s):
        out.append(s[a:])
    return out


def fmt_float(v, digits: int = 6) -> str:
    """``ofstream << float`` (default flags) at `digits` significant digits."""
    return "%.*g" % (digits, float(np.float32(v)))


def buffer_line(name: str, type_name: str, items) -> str:
    return f"{name}|{type_name}|" + "".join(f"{it};" for it in items) + "\n"


def f3(row, digits):
    return ",".join(fmt_float(c, digits) for c in row)


def mesh_text(x, tris, v, w, damping, gravity, digits=6) -> str:
    """PBDObject::SaveMeshTopol (PBD_cuda.cu:907-925) of a triangle mesh whose
    primList is (3t, 3) per triangle (IO::ReadTopolFromTxt, PBD_Basic.cuh:591)."""
    tris = np.asarray(tris, np.int64).reshape(-1, 3)
    out = buffer_line("Indices", T_INT, (str(int(i)) for i in tris.reshape(-1)))
    out += buffer_line("P", T_FLOAT3, (f3(r, digits) for r in np.asarray(x, np.float32).reshape(-1, 3)))
    out += buffer_line("primList", T_INT2, (f"{3 * t},3" for t in range(len(tris))))
    out += buffer_line("velBuffer", T_FLOAT3, (f3(r, digits) for r in np.asarray(v, np.float32).reshape(-1, 3)))
    out += buffer_line("massBuffer", T_FLOAT, (fmt_float(m, digits) for m in np.asarray(w, np.float32)))
    out += f"dampingRate|{T_FLOAT}|{fmt_float(damping, digits)}\n"
    out += f"gravity|{T_FLOAT3}|{f3(gravity, digits)}\n"
    return out


def constraint_text(pos0, p, prim_list, indices, rest, ne, w, stiffness, digits=6) -> str:
    """PBDObject::SaveConstraint (PBD_cuda.cu:627-640, 927-938) of prims laid
    out as the library builds them (ne distance prims, then bending prims),
    followed by one ANCHOR prim per weight-0 vertex in vertex order
    (ConstraintPBD::InitAnchorConstr, PBD_cuda.cu:122-146: one index, its
    position in restPosBuffer, stiffness -1), then this library's
    restAngleBuffer line when there are bending prims."""
    prim_list = [tuple(int(a) for a in r) for r in np.asarray(prim_list).reshape(-1, 2)]
    indices = [int(i) for i in np.asarray(indices).reshape(-1)]
    rest = np.asarray(rest, np.float32)
    pos0 = np.asarray(pos0, np.float32).reshape(-1, 3)
    nb = len(prim_list) - ne
    types = [DISTANCE] * ne + [BENDING] * nb
    stiff = [stiffness[0]] * ne + [stiffness[1]] * nb
    anchors = [i for i, wi in enumerate(np.asarray(w, np.float32)) if wi == 0.0]
    for a in anchors:
        prim_list.append((len(indices), 1))
        indices.append(a)
        types.append(ANCHOR)
        stiff.append(-1.0)
    out = buffer_line("Indices", T_INT, (str(i) for i in indices))
    out += buffer_line("P", T_FLOAT3, (f3(r, digits) for r in pos0))
    out += buffer_line("primList", T_INT2, (f"{s},{n}" for s, n in prim_list))
    out += buffer_line("prdPBuffer", T_FLOAT3, (f3(r, digits) for r in np.asarray(p, np.float32).reshape(-1, 3)))
    out += buffer_line("restPosBuffer", T_FLOAT3, (f3(pos0[a], digits) for a in anchors))
    out += buffer_line("restLengthBuffer", T_FLOAT, (fmt_float(r, digits) for r in rest[:ne]))
    out += buffer_line("stiffnessBuffer", T_FLOAT, (fmt_float(s, digits) for s in stiff))
    out += buffer_line("constraintType", T_INT, (str(t) for t in types))
    if nb:
        out += buffer_line("restAngleBuffer", T_FLOAT, (fmt_float(r, digits) for r in rest[ne:]))
    return out


def read_lines(text: str) -> dict[str, list[str]]:
    """name -> items, as the reference's readers see a file (a repeated name
    appends, as their push_back does)."""
    out: dict[str, list[str]] = {}
    for line in text.split("\n"):
        if line.endswith("\r"):
            line = line[:-1]
        parts = split(line, "|")
        if len(parts) != 3:
            continue
        out.setdefault(parts[0], []).extend(split(parts[2], ";"))
    return out


def _vec(items, k, conv):
    rows = []
    for it in items:
        c = split(it, ",")
        if len(c) != k:
            raise ValueError("wrong component count")
        rows.append([conv(x) for x in c])
    return rows


def _f32(s):
    return np.float32(float(s))


def read_cache(mesh_text_: str, constr_text_: str | None = None) -> dict:
    """The cloth IO::ReadTopolFromCache + IO::ReadConstraintFromCache describe,
    in the library's terms: start x, v, p, weights, triangles, distance pairs,
    bending quads, pins (ANCHOR prims), rest values in library prim order,
    stiffness per type, damping and gravity."""
    M = read_lines(mesh_text_)
    x = np.array(_vec(M["P"], 3, _f32), np.float32).reshape(-1, 3)
    nv = len(x)
    ind = [int(i) for i in M.get("Indices", [])]
    prims = _vec(M.get("primList", []), 2, int)
    if prims:
        tris = []
        for s, n in prims:
            if n != 3:
                raise ValueError("mesh prims must be triangles")
            tris += ind[s:s + 3]
    else:
        tris = ind
    out = {
        "x": x,
        "tris": np.array(tris, np.int32).reshape(-1, 3),
        "v": np.array(_vec(M["velBuffer"], 3, _f32), np.float32).reshape(-1, 3) if "velBuffer" in M
        else np.zeros_like(x),
        "w": np.array([_f32(m) for m in M["massBuffer"]], np.float32) if "massBuffer" in M
        else np.ones(nv, np.float32),
        "damping": _f32(M["dampingRate"][0]) if "dampingRate" in M else np.float32(1.0),
        "gravity": np.array(_vec(M["gravity"], 3, _f32)[0], np.float32) if "gravity" in M
        else np.array([0.0, -9.8, 0.0], np.float32),
        "pos0": x.copy(), "p": x.copy(), "pairs": None, "quads": None, "pins": [], "rest": None,
        "stiffness": [1.0, 1.0, 1.0],
    }
    if constr_text_ is None:
        return out
    K = read_lines(constr_text_)
    if "P" in K:
        out["pos0"] = np.array(_vec(K["P"], 3, _f32), np.float32).reshape(-1, 3)
    if "prdPBuffer" in K:
        out["p"] = np.array(_vec(K["prdPBuffer"], 3, _f32), np.float32).reshape(-1, 3)
    kind = [int(i) for i in K.get("Indices", [])]
    kprims = _vec(K.get("primList", []), 2, int)
    types = [int(t) for t in K.get("constraintType", [])]
    stiff = [_f32(s) for s in K.get("stiffnessBuffer", [])]
    rl = [_f32(r) for r in K.get("restLengthBuffer", [])]
    ra = [_f32(r) for r in K.get("restAngleBuffer", [])]
    pairs, quads, rest_d, rest_b, pins = [], [], [], [], []
    for (s, n), t, st in zip(kprims, types, stiff):
        q = kind[s:s + n]
        if t == DISTANCE:
            pairs.append(q)
            rest_d.append(rl[len(rest_d)])
            out["stiffness"][0] = float(st)
        elif t == BENDING:
            quads.append(q)
            rest_b.append(ra[len(rest_b)])
            out["stiffness"][1] = float(st)
        elif t == ANCHOR:
            if q[0] not in pins:
                pins.append(q[0])
        else:
            raise ValueError("unknown constraint type")
    out["pairs"] = np.array(pairs, np.int32).reshape(-1, 2)
    out["quads"] = np.array(quads, np.int32).reshape(-1, 4)
    out["pins"] = pins
    out["rest"] = np.array(rest_d + rest_b, np.float32)
    return out
