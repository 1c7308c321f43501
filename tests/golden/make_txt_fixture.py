"""Writes a small cloth in the reference's text formats (IO::ReadTopolFromTxt
and IO::ReadDistConstrFromTxt, PBD_Basic.cuh:532-688): topology file with
line 0 "P:x,y,z,..." and line 1 "Indices:i,j,k,...", constraint file with a
header line and then "label:i0,i1,..." lines (two of them here, to exercise
multi-line input).  Positions are a 6x5 grid in the XZ plane with seeded
jitter, printed with 9 significant digits so float32 values round-trip.

Run from the repository root: python tests/golden/make_txt_fixture.py
"""
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent
RX, RY = 6, 5


def main():
    rng = np.random.default_rng(20261202)
    pos = np.zeros((RY, RX, 3), np.float32)
    for j in range(RY):
        for i in range(RX):
            pos[j, i] = (i / (RX - 1), 0.0, j / (RY - 1))
    pos += rng.uniform(-1e-3, 1e-3, pos.shape).astype(np.float32)
    pos = pos.reshape(-1, 3).astype(np.float32)
    tris, pairs = [], []
    for j in range(RY - 1):
        for i in range(RX - 1):
            a = j * RX + i
            tris += [a, a + RX, a + RX + 1, a, a + RX + 1, a + 1]
    for j in range(RY):
        for i in range(RX):
            a = j * RX + i
            if i + 1 < RX:
                pairs += [a, a + 1]
            if j + 1 < RY:
                pairs += [a, a + RX]
            if j + 1 < RY and i > 0:
                pairs += [a, a + RX - 1]
    (OUT / "grid6x5_topol.txt").write_text(
        "P:" + ",".join(f"{v:.9g}" for v in pos.ravel()) + "\n" + "Indices:" + ",".join(map(str, tris)) + "\n")
    half = (len(pairs) // 4) * 2
    (OUT / "grid6x5_constr.txt").write_text(
        "header:ignored\n" + "Constr:" + ",".join(map(str, pairs[:half])) + "\n" + "Constr:" +
        ",".join(map(str, pairs[half:])) + "\n")
    np.savez(OUT / "grid6x5_expected.npz", pos=pos, tris=np.array(tris, np.int32).reshape(-1, 3),
             pairs=np.array(pairs, np.int32).reshape(-1, 2))


if __name__ == "__main__":
    main()
