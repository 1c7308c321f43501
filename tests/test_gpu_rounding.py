"""The distance projection takes its length and inverse length from short
forms of a correctly rounded square root and reciprocal (csrc/pbd_math.hpp:
the hardware estimates with one residual correction) wherever the squared
length lies in [2^-96, 2^96].  The solver's equality with the CPU path rests
on those forms returning sqrtf(x) and 1.0f / d exactly; tests/rounding_check
(built by build()) compares them on every float of their ranges on the GPU."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

CHECK = Path(__file__).resolve().parent / "rounding_check"


def test_short_sqrt_and_reciprocal_equal_ieee_on_every_float_of_their_range():
    print("rounding_check: every float of [2^-96, 2^97) and [2^-48, 2^49)", flush=True)
    assert CHECK.exists(), "tests/rounding_check is built by __graft_entry__.build()"
    r = subprocess.run([str(CHECK)], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all equal" in r.stdout
    lines = [l for l in r.stdout.splitlines() if l.startswith(("sqrt_rn", "rcp_rn"))]
    assert len(lines) == 3 and all(" 0 differ" in l for l in lines)
