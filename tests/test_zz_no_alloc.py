"""dbx_broadphase_pairs_noalloc: the stand-alone broadphase on caller-owned device
buffers with the library's fixed scratch, against dbx_broadphase_pairs (which
allocates its own buffers) on the same inputs, bit for bit, on two calls in a row,
the second with fewer proxies; plus the size limit.

The caller's device buffers come from the HIP runtime the library itself is
linked against (found in the process's memory map once the library is loaded),
so the test never brings a second runtime into the process."""
import ctypes as C

import numpy as np
import pytest

from dbox_amd import native

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip_runtime():
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64.so" in ln}
    assert len(paths) == 1, f"expected the library's one HIP runtime, found {paths}"
    rt = C.CDLL(paths.pop())
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return rt


class _Dev:
    """One device buffer holding a copy of a host array."""

    def __init__(self, rt, host):
        self.rt, self.host = rt, np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert rt.hipMalloc(C.byref(self.ptr), max(self.host.nbytes, 4)) == 0
        assert rt.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def read(self):
        out = np.empty_like(self.host)
        assert self.rt.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        assert self.rt.hipFree(self.ptr) == 0


def _fat(rng, n):
    lo = rng.uniform(-30, 30, (n, 2)).astype(np.float32)
    sz = rng.uniform(0.05, 2.5, (n, 2)).astype(np.float32)
    return np.concatenate([lo, lo + sz], 1).astype(np.float32)


def _noalloc(h, rt, fat, moved, cap):
    bufs = [_Dev(rt, fat.reshape(-1)), _Dev(rt, moved.astype(np.int32)),
            _Dev(rt, np.full(max(cap, 1) * 2, -7, np.int32)), _Dev(rt, np.full(1, -1, np.int32))]
    dfat, dmv, dout, dcnt = bufs
    st = h.L.dbx_broadphase_pairs_noalloc(h.h, dfat.ptr, dmv.ptr, len(fat), dout.ptr, cap, dcnt.ptr)
    assert st == native.DBX_OK, h.L.dbx_last_error(h.h)
    h.sync()
    n = int(dcnt.read()[0])
    pairs = dout.read().reshape(-1, 2)[:n] if n <= cap else None
    for b in bufs:
        b.free()
    return n, pairs


def test_noalloc_broadphase_matches_allocating_entry_point():
    print("case: no-alloc broadphase, 3000 then 700 proxies, seed 21", flush=True)
    rng = np.random.default_rng(21)
    h = native.StepWorlds()
    rt = _hip_runtime()
    for n in (3000, 700):
        fat = _fat(rng, n)
        moved = (rng.random(n) < 0.4).astype(np.uint8)
        want = native.broadphase_pairs(h, fat, moved)
        got_n, got = _noalloc(h, rt, fat, moved, 64 * n)
        assert got_n == len(want)
        assert got.tobytes() == want.astype(np.int32).tobytes()
    # a capacity below the count: the count is reported, no pair written
    fat = _fat(rng, 500)
    moved = np.ones(500, np.uint8)
    want = native.broadphase_pairs(h, fat, moved)
    got_n, got = _noalloc(h, rt, fat, moved, len(want) - 1)
    assert got_n == len(want) and got is None
    h.close()


def test_noalloc_broadphase_rejects_sizes_beyond_its_scratch():
    print("case: no-alloc broadphase size limit", flush=True)
    h = native.StepWorlds()
    n = 16385  # DBX_NOALLOC_MAX_PROXIES + 1: rejected before anything is launched
    dummy = C.c_void_p(256)
    st = h.L.dbx_broadphase_pairs_noalloc(h.h, dummy, dummy, n, dummy, 0, dummy)
    assert st == native.DBX_E_CAPACITY
    assert b"DBX_NOALLOC_MAX_PROXIES" in h.L.dbx_last_error(h.h)
    h.close()
