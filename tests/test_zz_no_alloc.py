"""pbd_sh_build_noalloc, the broadphase build that allocates nothing (device
arrays in and out, the sort's scratch in a static device array per GPU),
against the allocating entry point pbd_sh_build + pbd_sh_get on the same
input, bit for bit: sorted keys and ids, unsorted keys and cell ranges.
Two calls in a row on one GPU, the second on a smaller input, for a dense
grid (2 radix passes), a small hashed table and a 2^20 table (3 passes, the
sort ends in the scratch copy).  The test itself allocates the caller's
device arrays with the HIP runtime."""
import ctypes as C
import ctypes.util

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from positionbaseddynamicsold_amd import capi


def hip_runtime():
    for name in ("amdhip64", "libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        path = ctypes.util.find_library(name) if "/" not in name else name
        try:
            return C.CDLL(path or name)
        except OSError:
            continue
    raise RuntimeError("HIP runtime library not found")


class DeviceArrays:
    """The caller's device memory (hipMalloc / hipMemcpy / hipFree)."""

    def __init__(self):
        self.hip = hip_runtime()
        self.ptrs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 4))) == 0
        self.ptrs.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # H2D
        return p

    def filled(self, n, value=0x7F7F7F7F):
        return self.upload(np.full(max(n, 1), value, np.int32))

    def download(self, p, n):
        out = np.zeros(max(n, 1), np.int32)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0  # D2H
        return out[:n]

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def scene(n, seed):
    # a deformed n x n grid of triangles (two per quad)
    rng = np.random.default_rng(seed)
    s = np.linspace(0.0, 1.0, n, dtype=np.float64)
    gx, gz = np.meshgrid(s, s, indexing="xy")
    pos = np.stack([gx.ravel(), 0.05 * np.sin(6 * gx.ravel()) * np.cos(5 * gz.ravel()) +
                    0.01 * rng.standard_normal(n * n), gz.ravel()], 1).astype(np.float32)
    tris = []
    for y in range(n - 1):
        for x in range(n - 1):
            i = y * n + x
            tris += [(i, i + n, i + n + 1), (i, i + n + 1, i + 1)]
    return pos, np.array(tris, np.int32)


def noalloc(dev, desc, pos, tris, ncells):
    nv, nt = len(pos), len(tris)
    dpos, dtris = dev.upload(pos), dev.upload(tris)
    out = [dev.filled(nt) for _ in range(3)] + [dev.filled(ncells) for _ in range(2)]
    rc = capi.load().pbd_sh_build_noalloc(0, None, C.byref(desc), dpos, nv, dtris, nt, *out)
    assert rc == capi.PBD_OK, capi.load().pbd_last_error()
    return [dev.download(p, nt) for p in out[:3]] + [dev.download(p, ncells) for p in out[3:]]


@pytest.mark.parametrize("mode,bits", [(0, 0), (1, 10), (1, 20)])
def test_noalloc_broadphase_equals_allocating_build(mode, bits):
    print(f"noalloc broadphase mode={mode} bits={bits}", flush=True)
    desc = capi.hash_desc(mode, (0.5, 0.0, 0.5), (12, 4, 12), (0.1, 0.1, 0.1), bits)
    dev = DeviceArrays()
    try:
        with capi.Context() as ctx:
            for n, seed in ((90, 0), (23, 1)):  # 15,842 then 968 triangles
                pos, tris = scene(n, seed)
                ctx.sh_build(desc, pos, tris)
                want = ctx.sh_get()
                got = noalloc(dev, desc, pos, tris, len(want[3]))
                for name, a, b in zip(("sorted keys", "sorted ids", "keys", "cell start", "cell end"), got, want):
                    assert np.array_equal(a, b), f"{name} differ (n = {n})"
                assert (want[0][:-1] <= want[0][1:]).all()
                dev.free()
    finally:
        dev.free()


def test_noalloc_broadphase_limits_and_bad_indices():
    lib = capi.load()
    desc = capi.hash_desc(1, (0.5, 0.0, 0.5), (1, 1, 1), (0.1, 0.1, 0.1), 8)
    one = C.c_void_p(1)  # never dereferenced: the size check comes first
    rc = lib.pbd_sh_build_noalloc(0, None, C.byref(desc), one, 3, one, (1 << 20) + 1, one, one, one, one, one)
    assert rc == capi.PBD_ERR_LIMIT and b"maximum" in lib.pbd_last_error()
    # a triangle with an index outside [0, nv) is keyed -1; the others as usual
    pos, tris = scene(8, 2)
    bad = tris.copy()
    bad[5, 1] = len(pos) + 7
    dev = DeviceArrays()
    try:
        with capi.Context() as ctx:
            ctx.sh_build(desc, pos, tris)
            want = ctx.sh_get()
        got = noalloc(dev, desc, pos, bad, len(want[3]))
        raw = want[2].copy()
        raw[5] = -1
        assert np.array_equal(got[2], raw)
    finally:
        dev.free()
