"""The C-ABI library loads and exports every entry point include/pbd.h declares.
CPU only: no compute call is made."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        names |= set(re.findall(r"^PBD_API\s+[\w\s\*]+?\b(pbd_\w+)\s*\(", h.read_text(), re.M))
    return names


def test_header_declares_entry_points():
    names = declared()
    assert {"pbd_create", "pbd_init", "pbd_step", "pbd_cloth_get_colors"} <= names


def test_library_exports_every_declared_symbol():
    from positionbaseddynamicsold_amd import capi
    lib = capi.load()
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes signature table covers every declared function
    assert set(capi.SIGNATURES) | {"pbd_last_error"} >= declared()


def test_no_internal_symbols_exported():
    import subprocess
    lib = ROOT / "positionbaseddynamicsold_amd" / "libpbd.so"
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib)], capture_output=True, text=True).stdout
    text = [l.split()[-1] for l in out.splitlines() if " T " in l]
    assert text and all(t.startswith("pbd_") for t in text), text


def test_last_error_is_empty_string_initially():
    from positionbaseddynamicsold_amd import capi
    assert capi.load().pbd_last_error() == b""
