"""kc_decompose_noalloc against kc_decompose, through the C ABI.

The caller owns the device memory: the CSR and the core buffer are
allocated here with the HIP runtime libkcore itself links
(libamdhip64.so.7), and the entry point carves its work buffers from the
library's static device array.  Two calls in a row, the second on a smaller
graph (the array keeps the first call's contents), each bit-exact against
kc_decompose and the oracle's Batagelj-Zaversnik cores; sizes beyond the
array are refused before anything runs on the device."""
import ctypes

import numpy as np
import pytest

import oracle
import pico_amd as pa
from pico_amd import kcore

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return h


def _lib():
    L = kcore.lib()
    f = L.kc_decompose_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(kcore.Opts), ctypes.c_void_p,
                  ctypes.POINTER(kcore.Stats)]
    f.restype = ctypes.c_int
    return L


class _Dev:
    """Device copies of a CSR (u32 row offsets) and a core buffer."""

    def __init__(self, hip, rowptr, col):
        self.hip = hip
        self.V = len(rowptr) - 1
        self.nnz = int(rowptr[-1])
        self.ptrs = []
        self.rp = self._put(np.ascontiguousarray(rowptr, np.uint32))
        self.col = self._put(np.ascontiguousarray(col, np.uint32)) if self.nnz else None
        self.core = self._alloc(max(self.V, 1) * 4)

    def _alloc(self, n):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), n) == 0
        self.ptrs.append(p)
        return p

    def _put(self, a):
        p = self._alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0
        return p

    def cores(self):
        out = np.zeros(max(self.V, 1), np.uint32)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.core, out.nbytes,
                                  D2H) == 0
        return out[:self.V]

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _noalloc(L, ctx, d, method=3, **opts):
    st = kcore.Stats()
    o = kcore.make_opts(**opts)
    s = L.kc_decompose_noalloc(ctx._h, d.V, d.nnz, d.rp, d.col, method, ctypes.byref(o), d.core,
                               ctypes.byref(st))
    return s, st.as_dict()


def test_noalloc_equals_decompose_twice_in_a_row():
    hip, L = _hip(), _lib()
    ctx = pa.Context(0)
    try:
        # the larger graph first, then a smaller one over the same array
        for scale, ef, seed, method in ((17, 16, 3, 3), (11, 8, 4, 2)):
            hg = pa.HostGraph.rmat(scale, ef, seed=seed)
            want = oracle.bz(hg.rowptr, hg.col)
            dg = ctx.upload(hg)
            ref, rst = dg.decompose(pa.kcore.METHOD_NAMES[method])
            dg.close()
            np.testing.assert_array_equal(ref, want)
            d = _Dev(hip, hg.rowptr, hg.col)
            try:
                s, st = _noalloc(L, ctx, d, method)
                assert s == 0, L.kc_last_error()
                got = d.cores()
            finally:
                d.free()
            print("[noalloc] scale %d: %.3f ms, kmax %d" % (scale, st["device_ms"], st["kmax"]),
                  flush=True)
            np.testing.assert_array_equal(got, ref)
            assert st["kmax"] == rst["kmax"] == int(want.max())
            assert st["iterations"] == rst["iterations"] == int(want.max()) + 1
    finally:
        ctx.close()


def test_noalloc_small_graphs_and_launch_shapes():
    hip, L = _hip(), _lib()
    ctx = pa.Context(0)
    try:
        cases = [pa.HostGraph.er(10000, 50000, seed=0), pa.HostGraph.er(3, 0, seed=1),
                 pa.HostGraph.chunglu(1 << 14, beta=2.1, mean_deg=20.0, cap=3000.0, seed=2)]
        for hg in cases:
            want = oracle.bz(hg.rowptr, hg.col)
            for block in (256, 1024):
                d = _Dev(hip, hg.rowptr, hg.col)
                try:
                    s, _ = _noalloc(L, ctx, d, 2, block=block)
                    assert s == 0, L.kc_last_error()
                    np.testing.assert_array_equal(d.cores(), want)
                finally:
                    d.free()
    finally:
        ctx.close()


def test_noalloc_refuses_what_does_not_fit():
    hip, L = _hip(), _lib()
    ctx = pa.Context(0)
    d = _Dev(hip, np.zeros(2, np.uint32), np.zeros(0, np.uint32))
    try:
        # sizes beyond the static array: refused on the host, nothing launched
        st = kcore.Stats()
        s = L.kc_decompose_noalloc(ctx._h, 50_000_000, 0, d.rp, None, 3, None, d.core,
                                   ctypes.byref(st))
        assert s == 2 and b"static array" in L.kc_last_error()
        s = L.kc_decompose_noalloc(ctx._h, 1, 0, d.rp, None, 4, None, d.core, ctypes.byref(st))
        assert s == 1  # not a bucket peel
        s = L.kc_decompose_noalloc(ctx._h, 1, 0, None, None, 3, None, d.core, ctypes.byref(st))
        assert s == 1  # null rows
    finally:
        d.free()
        ctx.close()
