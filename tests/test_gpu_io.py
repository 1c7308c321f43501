"""Reference text formats through pbd_cloth_load_txt (IO::ReadTopolFromTxt
and IO::ReadDistConstrFromTxt, PBD_Basic.cuh:532-688) against the fixture's
own arrays (tests/golden, written by make_txt_fixture.py) and the oracle:
parsed positions, triangles, distance prims in file order, rest lengths by
the reference's powf expression, and the stepped result; malformed and
missing files are reported as PBD_ERR_IO."""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from positionbaseddynamicsold_amd import capi

GOLD = Path(__file__).resolve().parent / "golden"


def test_load_txt_matches_fixture(orc):
    exp = np.load(GOLD / "grid6x5_expected.npz")
    pos, tris, pairs = exp["pos"], exp["tris"], exp["pairs"]
    with capi.Context() as ctx:
        cid = ctx.load_txt(str(GOLD / "grid6x5_topol.txt"), str(GOLD / "grid6x5_constr.txt"))
        ref = ctx.add_mesh(pos, tris)
        ctx.set_distance_pairs(ref, pairs)
        for c in (cid, ref):
            ctx.set_pins(c, [0, 5])
        ctx.init()
        inf = ctx.info(cid)
        assert (inf.nv, inf.nt, inf.ne, inf.nb) == (30, len(tris), len(pairs), 0)
        assert np.array_equal(ctx.get(cid), pos)
        pl, idx, rest = ctx.prims(cid)
        opl, oidx, orest = orc.build_prims(pos, pairs)
        assert np.array_equal(pl, opl) and np.array_equal(idx, oidx)
        assert np.array_equal(rest, orest)
        ctx.set_solver(dt=1.0 / 600.0, iterations=20)
        ctx.step(30)
        oc = orc.OracleCloth(pos, tris, pairs, pins=[0, 5])
        prm = oc.params()
        for _ in range(30):
            oc.substep(prm, cap=20)
        assert np.array_equal(ctx.get(cid), oc.x)
        assert np.array_equal(ctx.get(cid), ctx.get(ref))


def test_load_txt_without_constraint_file_uses_mesh_edges():
    with capi.Context() as ctx:
        cid = ctx.load_txt(str(GOLD / "grid6x5_topol.txt"))
        ctx.init()
        exp = np.load(GOLD / "grid6x5_expected.npz")
        edges = {tuple(sorted(e)) for t in exp["tris"] for e in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))}
        assert ctx.info(cid).ne == len(edges)


def test_load_txt_errors(tmp_path):
    bad = tmp_path / "bad.txt"
    bad.write_text("P:0,0,0,1,x,0\nIndices:0,1,2\n")
    odd = tmp_path / "odd.txt"
    odd.write_text("P:0,0,0,1,0\n")
    with capi.Context() as ctx:
        for path in (tmp_path / "missing.txt", bad, odd):
            with pytest.raises(capi.PbdError) as e:
                ctx.load_txt(str(path))
            assert e.value.code == capi.PBD_ERR_IO, path
        with pytest.raises(capi.PbdError) as e:
            ctx.load_txt(str(GOLD / "grid6x5_topol.txt"), str(tmp_path / "missing_constr.txt"))
        assert e.value.code == capi.PBD_ERR_IO
