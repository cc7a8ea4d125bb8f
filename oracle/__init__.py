"""CPU restatement of the PICO k-core methods -- TEST INFRASTRUCTURE ONLY.

This package is the checker.  Only tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import it; the product path
(``pico_amd``) never does.  The C sources (``kc_oracle.c``) cite the reference
file:line each function restates; ``_ref/`` (built by ``make`` here when the
reference tree is mounted) holds the reference's own host readers.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "liboracle.so")
REF_COO = os.path.join(HERE, "_ref", "ref_coo_dump")
REF_CSR = os.path.join(HERE, "_ref", "ref_csr_dump")

_lib = None


def build(ref=True):
    """Compile liboracle.so (and _ref/ when the reference tree is present)."""
    target = "all" if ref else "lib/liboracle.so"
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(ref=False)
        L = ctypes.CDLL(LIB)
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
        c_u64p = ctypes.POINTER(ctypes.c_uint64)
        L.orc_bz.argtypes = [ctypes.c_uint32, u64p, u32p, u32p]
        L.orc_bz.restype = ctypes.c_uint32
        L.orc_gpp.argtypes = [ctypes.c_uint32, u64p, u32p, ctypes.c_int, u32p, c_u64p]
        L.orc_peelone.argtypes = [ctypes.c_uint32, u64p, u32p, ctypes.c_uint32, u32p, c_u64p]
        L.orc_bucket_peel.argtypes = [ctypes.c_uint32, u64p, u32p, u32p, c_u64p]
        L.orc_bucket_peel.restype = ctypes.c_uint32
        L.orc_histocore.argtypes = [ctypes.c_uint32, u64p, u32p, u32p, c_u64p, ctypes.c_void_p]
        L.orc_cntcore.argtypes = [ctypes.c_uint32, u64p, u32p, u32p, c_u64p]
        L.orc_nbrcore.argtypes = [ctypes.c_uint32, u64p, u32p, u32p, c_u64p]
        L.orc_hindex_jacobi.argtypes = [ctypes.c_uint32, u64p, u32p, u32p, c_u64p]
        L.orc_bruteforce.argtypes = [ctypes.c_uint32, u64p, u32p, u32p]
        L.orc_check_invariants.argtypes = [ctypes.c_uint32, u64p, u32p, u32p]
        L.orc_check_invariants.restype = ctypes.c_uint64
        _lib = L
    return _lib


def _csr(rowptr, col):
    rp = np.ascontiguousarray(rowptr, dtype=np.uint64)
    cl = np.ascontiguousarray(col, dtype=np.uint32)
    if cl.size == 0:
        cl = np.zeros(1, np.uint32)
    return rp, cl, len(rp) - 1


def bz(rowptr, col):
    """Batagelj-Zaversnik core numbers (uint32[V])."""
    rp, cl, V = _csr(rowptr, col)
    core = np.zeros(max(V, 1), np.uint32)
    lib().orc_bz(V, rp, cl, core)
    return core[:V]


def _iter_call(fn, rowptr, col, *extra):
    rp, cl, V = _csr(rowptr, col)
    core = np.zeros(max(V, 1), np.uint32)
    it = ctypes.c_uint64(0)
    fn(V, rp, cl, *extra, core, ctypes.byref(it)) if extra else fn(V, rp, cl, core, ctypes.byref(it))
    return core[:V], int(it.value)


def gpp(rowptr, col, compat=True):
    """(core, iterations) of gppeel; compat keeps its isolated-vertex=1 quirk."""
    return _iter_call(lib().orc_gpp, rowptr, col, int(bool(compat)))


def peelone(rowptr, col, part_size=64):
    return _iter_call(lib().orc_peelone, rowptr, col, int(part_size))


def bucket_peel(rowptr, col):
    """(core, levels) of peel_dyn / peelone_dyn; kmax = max(core)."""
    rp, cl, V = _csr(rowptr, col)
    core = np.zeros(max(V, 1), np.uint32)
    lv = ctypes.c_uint64(0)
    lib().orc_bucket_peel(V, rp, cl, core, ctypes.byref(lv))
    return core[:V], int(lv.value)


def histocore(rowptr, col, want_hist=False):
    rp, cl, V = _csr(rowptr, col)
    core = np.zeros(max(V, 1), np.uint32)
    it = ctypes.c_uint64(0)
    hist = np.zeros(int(rp[-1]) + V + 1, np.uint32) if want_hist else None
    lib().orc_histocore(V, rp, cl, core, ctypes.byref(it),
                        hist.ctypes.data if want_hist else None)
    if want_hist:
        return core[:V], int(it.value), hist[: int(rp[-1]) + V]
    return core[:V], int(it.value)


def cntcore(rowptr, col):
    return _iter_call(lib().orc_cntcore, rowptr, col)


def nbrcore(rowptr, col):
    return _iter_call(lib().orc_nbrcore, rowptr, col)


def hindex_jacobi(rowptr, col):
    return _iter_call(lib().orc_hindex_jacobi, rowptr, col)


def bruteforce(rowptr, col):
    rp, cl, V = _csr(rowptr, col)
    core = np.zeros(max(V, 1), np.uint32)
    lib().orc_bruteforce(V, rp, cl, core)
    return core[:V]


def check_invariants(rowptr, col, core):
    """0 when core is an h-index fixed point bounded by degree, else 1+vertex."""
    rp, cl, V = _csr(rowptr, col)
    c = np.ascontiguousarray(core, dtype=np.uint32)
    if c.size == 0:
        c = np.zeros(1, np.uint32)
    return int(lib().orc_check_invariants(V, rp, cl, c))


def ref_available():
    return os.path.exists(REF_COO) and os.path.exists(REF_CSR)


def ref_read_coo(path, workdir):
    """Run the reference's COO reader + MakeUGraph; return its layout."""
    out = os.path.join(workdir, "coo.bin")
    subprocess.run([REF_COO, os.path.abspath(path), out], cwd=workdir, check=True,
                   stdout=subprocess.DEVNULL)
    w = np.fromfile(out, dtype=np.uint32)
    n, e, plo, phi, zero = (int(x) for x in w[:5])
    P = plo | (phi << 32)
    o = 5
    node_ptr = w[o:o + n]
    o += n
    edge_list = w[o:o + 2 * e + n]
    o += 2 * e + n
    parts = w[o:o + 2 * P].reshape(-1, 2) if P else np.zeros((0, 2), np.uint32)
    return dict(num_nodes=n, num_edges=e, has_zero=bool(zero), node_pointer=node_ptr,
                edge_list=edge_list, parts=parts)


def ref_read_csr(path, workdir):
    """Run the reference's deduplicating CSR reader (peelone_dyn input)."""
    out = os.path.join(workdir, "csr.bin")
    subprocess.run([REF_CSR, os.path.abspath(path), out], cwd=workdir, check=True,
                   stdout=subprocess.DEVNULL)
    w = np.fromfile(out, dtype=np.uint32)
    V, E = int(w[0]), int(w[1])
    o = 2
    deg = w[o:o + V]
    o += V
    off = w[o:o + V + 1]
    o += V + 1
    nb = w[o:o + E]
    return dict(V=V, E=E, degrees=deg, offsets=off, neighbors=nb)
