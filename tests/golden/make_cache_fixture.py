"""Writes grid3x2_meshtopol.cache and grid3x2_constraint.cache: a 3 x 2 grid
cloth part way through a simulation, in the text the reference's writers
produce (PBDObject::SaveMeshTopol / SaveConstraint, PBD_cuda.cu:907-938), typed
out by hand from those writers rather than produced by any code under test:

- Windows line ends (the reference is a Visual Studio project);
- MSVC type names ("struct float3");
- 6 significant digits (the stream default);
- an empty mesh primList line (a grid built by initMeshTopolIndices,
  PBD_cuda.cu:855-880, fills no mesh primList);
- masses all 1 with the two pinned corners held by ANCHOR prims
  (InitAnchorConstr, stiffness -1, PBD_cuda.cu:122-146);
- a restPosBuffer holding four positions (initPosition pushes the two corners,
  PBD_cuda.cu:850-851, and InitAnchorConstr pushes them again);
- dampingRate and gravity written by IO::SaveData without a trailing ';'.
"""
from pathlib import Path

HERE = Path(__file__).resolve().parent

MESH = [
    "Indices|int|0;1;3;1;4;3;1;2;4;2;5;4;",
    "P|struct float3|0,0,0;0.5,-0.0012,0.01;1,0,0;0.0101,-0.0503,0.4991;0.5,-0.0501,0.5;0.9899,-0.0503,0.4991;",
    "primList|struct int2|",
    "velBuffer|struct float3|0,0,0;0,-0.72,0.06;0,0,0;0.0606,-0.3018,-0.0054;0,-0.3006,0;-0.0606,-0.3018,-0.0054;",
    "massBuffer|float|1;1;1;1;1;1;",
    "dampingRate|float|0.9",
    "gravity|struct float3|0,-10,0",
]
CONSTRAINT = [
    "Indices|int|0;1;0;3;1;2;1;4;2;5;3;4;4;5;0;2;",
    "P|struct float3|0,0,0;0.5,0,0;1,0,0;0,0,0.5;0.5,0,0.5;1,0,0.5;",
    "primList|struct int2|0,2;2,2;4,2;6,2;8,2;10,2;12,2;14,1;15,1;",
    "prdPBuffer|struct float3|0,0,0;0.5,-0.0013,0.0101;1,0,0;0.0102,-0.0508,0.499;0.5,-0.0506,0.5;"
    "0.9898,-0.0508,0.499;",
    "restPosBuffer|struct float3|0,0,0;1,0,0;0,0,0;1,0,0;",
    "restLengthBuffer|float|0.5;0.5;0.5;0.5;0.5;0.5;0.5;",
    "stiffnessBuffer|float|1;1;1;1;1;1;1;-1;-1;",
    "constraintType|int|1;1;1;1;1;1;1;256;256;",
]


def main():
    for name, lines in (("grid3x2_meshtopol.cache", MESH), ("grid3x2_constraint.cache", CONSTRAINT)):
        (HERE / name).write_bytes("".join(line + "\r\n" for line in lines).encode())


if __name__ == "__main__":
    main()
