"""Python view of libkcore's C ABI (include/kcore.h).

The methods keep the reference's names and meanings:

    gpp, peelone, peel_dyn, peelone_dyn      peel family  (peel/*.cu, peel/src/*.cc)
    HistoCore, CntCore, NbrCore              h-index family (index2core/*.cu)

Every call goes to the HIP library; there is no CPU fallback.  A missing or
unloadable libkcore.so raises ``KCoreError`` at import of this module's first
use.
"""
import atexit
import ctypes
import os
import sys
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libkcore.so")

METHODS = {
    "gpp": 0, "peelone": 1, "peel_dyn": 2, "peelone_dyn": 3,
    "histocore": 4, "cntcore": 5, "nbrcore": 6,
}
METHOD_NAMES = {v: k for k, v in METHODS.items()}
STATUS = {0: "OK", 1: "EINVAL", 2: "ENOMEM", 3: "EDEVICE", 4: "EOVERFLOW", 5: "ETIMEOUT",
          6: "EIO"}


class KCoreError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS.get(status, status), msg))
        self.status = status


class Opts(ctypes.Structure):
    _fields_ = [("block", ctypes.c_uint32), ("grid", ctypes.c_uint32),
                ("hub_split", ctypes.c_uint32), ("queue_cap", ctypes.c_uint32),
                ("gpp_compat", ctypes.c_uint32), ("part_size", ctypes.c_uint32),
                ("local_cap", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("agg_div", ctypes.c_uint32), ("compact_div", ctypes.c_uint32),
                ("stage_grid", ctypes.c_uint32), ("pull_div", ctypes.c_uint32)]


class Stats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint64), ("levels", ctypes.c_uint64),
                ("levels_run", ctypes.c_uint64), ("phases", ctypes.c_uint64),
                ("edges_walked", ctypes.c_uint64), ("kmax", ctypes.c_uint32),
                ("grid", ctypes.c_uint32), ("block", ctypes.c_uint32),
                ("stages", ctypes.c_uint32), ("device_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("changed", ctypes.c_uint64),
                ("prep_ms", ctypes.c_double), ("changed_degree", ctypes.c_uint64),
                ("atomics", ctypes.c_uint64), ("rebuild_rounds", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_
                if not k.startswith("reserved")}


class GraphInfo(ctypes.Structure):
    _fields_ = [("V", ctypes.c_uint32), ("dmax", ctypes.c_uint32), ("nnz", ctypes.c_uint64),
                ("edges_read", ctypes.c_uint64), ("has_zero_id", ctypes.c_uint32),
                ("multigraph", ctypes.c_uint32), ("rowptr", ctypes.POINTER(ctypes.c_uint64)),
                ("col", ctypes.POINTER(ctypes.c_uint32))]


class ReadStats(ctypes.Structure):
    _fields_ = [("V", ctypes.c_uint32), ("has_zero_id", ctypes.c_uint32),
                ("multigraph", ctypes.c_uint32), ("threads", ctypes.c_uint32),
                ("nnz", ctypes.c_uint64), ("edges_read", ctypes.c_uint64),
                ("file_bytes", ctypes.c_uint64), ("parse_ms", ctypes.c_double),
                ("build_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("reserved", ctypes.c_uint64 * 4)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class RefGraph(ctypes.Structure):
    """Mirror of the peelone_dyn Graph's data members (peel/inc/graph.h:7-23)."""
    _fields_ = [("V", ctypes.c_uint32), ("E", ctypes.c_uint32),
                ("neighbors", ctypes.POINTER(ctypes.c_uint32)),
                ("neighbors_offset", ctypes.POINTER(ctypes.c_uint32)),
                ("degrees", ctypes.POINTER(ctypes.c_uint32)),
                ("kmax", ctypes.c_uint32), ("dmax", ctypes.c_uint32)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("G", ctypes.c_uint32), ("reserved0", ctypes.c_uint32),
                ("vertices", ctypes.c_uint64), ("entries", ctypes.c_uint64),
                ("vtx_off", ctypes.POINTER(ctypes.c_uint64)),
                ("edge_off", ctypes.POINTER(ctypes.c_uint64)),
                ("rowptr", ctypes.POINTER(ctypes.c_uint32)),
                ("col", ctypes.POINTER(ctypes.c_uint32)),
                ("scale", ctypes.POINTER(ctypes.c_uint32)),
                ("m_gen", ctypes.POINTER(ctypes.c_uint64))]


class BatchOpts(ctypes.Structure):
    _fields_ = [("block", ctypes.c_uint32), ("per_cu", ctypes.c_uint32),
                ("lds_vertices", ctypes.c_uint32), ("split", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 4)]


class BatchStats(ctypes.Structure):
    _fields_ = [("graphs", ctypes.c_uint64), ("vertices", ctypes.c_uint64),
                ("entries", ctypes.c_uint64), ("levels_run", ctypes.c_uint64),
                ("kmax", ctypes.c_uint32), ("block", ctypes.c_uint32),
                ("grid", ctypes.c_uint32), ("launches", ctypes.c_uint32),
                ("device_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("reserved", ctypes.c_uint64 * 4)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_
                if not k.startswith("reserved")}


class UpdateStats(ctypes.Structure):
    _fields_ = [("requested", ctypes.c_uint64), ("applied", ctypes.c_uint64),
                ("skipped", ctypes.c_uint64), ("candidates", ctypes.c_uint64),
                ("evaluations", ctypes.c_uint64), ("changed", ctypes.c_uint64),
                ("matchings", ctypes.c_uint32), ("bfs_levels", ctypes.c_uint32),
                ("rounds", ctypes.c_uint32), ("grows", ctypes.c_uint32),
                ("device_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("reserved", ctypes.c_uint64 * 4)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_
                if not k.startswith("reserved")}


class ServerOpts(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_uint32), ("max_batch", ctypes.c_uint32),
                ("fuse_max_scale", ctypes.c_uint32), ("reserved0", ctypes.c_uint32),
                ("hbm_budget", ctypes.c_uint64), ("reserved", ctypes.c_uint64 * 4)]


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("graph", ctypes.c_uint32),
                ("op", ctypes.c_uint32), ("scale", ctypes.c_uint32),
                ("edge_factor", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("n", ctypes.c_uint64), ("u", ctypes.c_void_p), ("v", ctypes.c_void_p),
                ("core_out", ctypes.c_void_p), ("tag", ctypes.c_uint64),
                ("reserved", ctypes.c_uint64 * 4)]


class Completion(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("tag", ctypes.c_uint64),
                ("status", ctypes.c_int32), ("kind", ctypes.c_uint32),
                ("V", ctypes.c_uint32), ("kmax", ctypes.c_uint32),
                ("fused", ctypes.c_uint32), ("reserved0", ctypes.c_uint32),
                ("submit_ms", ctypes.c_double), ("start_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double), ("checksum", ctypes.c_uint64),
                ("applied", ctypes.c_uint64), ("reserved", ctypes.c_uint64 * 4)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_
                if not k.startswith("reserved")}


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in (
        "submitted", "completed", "failed", "jobs", "fused_jobs", "fused_requests",
        "admission_waits", "max_busy", "max_inflight_bytes", "hbm_budget")] + \
        [("reserved", ctypes.c_uint64 * 4)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_
                if not k.startswith("reserved")}


_lib = None
_vp = ctypes.c_void_p
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def lib():
    """Load libkcore.so (built by __graft_entry__.build / pico_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise KCoreError(3, "libkcore.so is not built (%s); run pico_amd.build" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    S = ctypes.c_int
    sig = {
        "kc_last_error": ([], ctypes.c_char_p),
        "kc_version": ([], ctypes.c_char_p),
        "kc_hg_read_coo": ([ctypes.c_char_p, ctypes.POINTER(_vp)], S),
        "kc_hg_save": ([_vp, ctypes.c_char_p], S),
        "kc_hg_load": ([ctypes.c_char_p, ctypes.POINTER(_vp)], S),
        "kc_hg_read_cached": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                               ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_int)], S),
        "kc_hg_read_dedup": ([ctypes.c_char_p, ctypes.POINTER(_vp)], S),
        "kc_hg_from_pairs": ([ctypes.c_uint32, ctypes.c_uint64, _vp, _vp, ctypes.POINTER(_vp)], S),
        "kc_hg_from_csr": ([ctypes.c_uint32, _vp, _vp, ctypes.POINTER(_vp)], S),
        "kc_hg_gen_er": ([ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(_vp)], S),
        "kc_hg_gen_rmat": ([ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                            ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(_vp)], S),
        "kc_hg_gen_chunglu": ([ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                               ctypes.c_uint64, ctypes.POINTER(_vp)], S),
        "kc_hg_info": ([_vp, ctypes.POINTER(GraphInfo)], S),
        "kc_hg_vgraph": ([_vp, ctypes.c_uint32, _vp, _vp, _vp, _u64p], S),
        "kc_hg_write_edgelist": ([_vp, ctypes.c_char_p], S),
        "kc_hg_free": ([_vp], None),
        "kc_write_cores": ([ctypes.c_char_p, _vp, ctypes.c_uint32], S),
        "kc_ctx_create": ([ctypes.c_int, ctypes.POINTER(_vp)], S),
        "kc_ctx_destroy": ([_vp], None),
        "kc_ctx_info": ([_vp, _u32p, _u64p], S),
        "kc_graph_upload": ([_vp, ctypes.c_uint32, _vp, _vp, ctypes.c_uint32, ctypes.POINTER(_vp)], S),
        "kc_graph_upload_hg": ([_vp, _vp, ctypes.c_uint32, ctypes.POINTER(_vp)], S),
        "kc_graph_free": ([_vp], None),
        "kc_graph_device_bytes": ([_vp], ctypes.c_uint64),
        "kc_decompose": ([_vp, ctypes.c_int, ctypes.POINTER(Opts), _vp, ctypes.POINTER(Stats)], S),
        "kc_graph_core_device": ([_vp], _vp),
        "kc_peel_dyn": ([ctypes.POINTER(RefGraph), _vp], ctypes.c_int),
        "kc_peelone_dyn": ([ctypes.POINTER(RefGraph), _vp], ctypes.c_int),
        "kc_vgraph_decompose": ([ctypes.c_uint32, ctypes.c_uint64, _vp, _vp, ctypes.c_int,
                                 ctypes.POINTER(Opts), _vp, ctypes.POINTER(Stats)], S),
        "kc_hb_gen_rmat": ([ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double,
                            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                            ctypes.c_uint64, ctypes.POINTER(_vp)], S),
        "kc_hb_from_graphs": ([ctypes.c_uint32, ctypes.POINTER(_vp), ctypes.POINTER(_vp)], S),
        "kc_hb_info": ([_vp, ctypes.POINTER(BatchInfo)], S),
        "kc_hb_graph": ([_vp, ctypes.c_uint32, ctypes.POINTER(_vp)], S),
        "kc_hb_free": ([_vp], None),
        "kc_batch_upload": ([_vp, ctypes.c_uint32, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp)], S),
        "kc_batch_upload_hb": ([_vp, _vp, ctypes.POINTER(_vp)], S),
        "kc_batch_run": ([_vp, ctypes.POINTER(BatchOpts), _vp, _vp,
                          ctypes.POINTER(BatchStats)], S),
        "kc_batch_run_plain": ([_vp, ctypes.c_int, ctypes.POINTER(Opts), _vp, _vp,
                                ctypes.POINTER(BatchStats)], S),
        "kc_batch_autotune": ([_vp, ctypes.c_uint32, ctypes.POINTER(BatchOpts),
                               ctypes.POINTER(ctypes.c_double)], S),
        "kc_batch_decompose": ([_vp, ctypes.c_uint32, _vp, _vp, _vp, _vp,
                                ctypes.POINTER(BatchOpts), _vp, ctypes.POINTER(BatchStats)], S),
        "kc_batch_device_bytes": ([_vp], ctypes.c_uint64),
        "kc_batch_free": ([_vp], None),
        "kc_autotune": ([_vp, ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(Opts),
                         ctypes.POINTER(ctypes.c_double)], S),
        "kc_graph_gen_rmat": ([_vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double,
                               ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                               ctypes.c_uint32, ctypes.POINTER(_vp)], S),
        "kc_graph_gen_chunglu": ([_vp, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32,
                                  ctypes.POINTER(_vp)], S),
        "kc_graph_download": ([_vp, ctypes.POINTER(_vp)], S),
        "kc_graph_read": ([_vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_uint32,
                           ctypes.POINTER(_vp), ctypes.POINTER(ReadStats)], S),
        "kc_batch_gen_rmat": ([_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                               ctypes.c_double, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                               ctypes.POINTER(_vp)], S),
        "kc_batch_download": ([_vp, ctypes.POINTER(_vp)], S),
        "kc_batch_dims": ([_vp, _u32p, _u64p, _u64p], S),
        "kc_update": ([_vp, ctypes.c_int, _vp, _vp, ctypes.c_uint64, _vp,
                       ctypes.POINTER(UpdateStats)], S),
        "kc_graph_dims": ([_vp, _u32p, _u64p], S),
        "kc_graph_degree_stats": ([_vp, _vp], S),
        "kc_update_prepare": ([_vp], S),
        "kc_kcore_subgraph": ([_vp, ctypes.c_uint32, ctypes.POINTER(_vp), _vp, _u32p], S),
        "kc_server_create": ([ctypes.c_int, ctypes.POINTER(ServerOpts), ctypes.POINTER(_vp)], S),
        "kc_server_gen_rmat_graph": ([_vp, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64,
                                      _u32p], S),
        "kc_server_graph_download": ([_vp, ctypes.c_uint32, ctypes.POINTER(_vp)], S),
        "kc_server_submit": ([_vp, ctypes.POINTER(Request), _u64p], S),
        "kc_server_poll": ([_vp, ctypes.POINTER(Completion), ctypes.c_uint32, _u32p,
                            ctypes.c_int], S),
        "kc_server_drain": ([_vp], S),
        "kc_server_now_ms": ([_vp], ctypes.c_double),
        "kc_server_get_stats": ([_vp, ctypes.POINTER(ServerStats)], S),
        "kc_server_destroy": ([_vp], None),
        "kc_core_checksum": ([_vp, ctypes.c_uint32], ctypes.c_uint64),
        "kc_probe_copy": ([_vp, ctypes.c_uint64, ctypes.c_uint32,
                           ctypes.POINTER(ctypes.c_double)], S),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


# Every live handle object, closed in dependency order when the interpreter
# exits (servers, then device objects, then host objects, then contexts),
# while the HIP runtime is still up; finalizers run during interpreter
# shutdown make no library calls at all.
_live = weakref.WeakSet()
_CLOSE_ORDER = ("Server", "DeviceBatch", "DeviceGraph", "HostBatch", "HostGraph", "Context")


def _register(obj):
    _live.add(obj)


def _close_all():
    objs = list(_live)
    for name in _CLOSE_ORDER:
        for o in objs:
            if type(o).__name__ == name:
                try:
                    o.close()
                except Exception:
                    pass


atexit.register(_close_all)


def _finalize(obj):
    if sys.is_finalizing():
        return
    try:
        obj.close()
    except Exception:
        pass


def _check(status):
    if status != 0:
        raise KCoreError(status, lib().kc_last_error().decode(errors="replace"))


def _ptr(a):
    return a.ctypes.data_as(_vp) if a is not None and a.size else None


def make_opts(**kw):
    o = Opts()
    for k, v in kw.items():
        if v is not None:
            setattr(o, k, int(v))
    return o


# ------------------------------------------------------------- host graphs
class HostGraph:
    """Host CSR owned by libkcore (kc_host_graph)."""

    def __init__(self, handle):
        self._h = handle
        _register(self)
        info = GraphInfo()
        _check(lib().kc_hg_info(self._h, ctypes.byref(info)))
        self.V = int(info.V)
        self.nnz = int(info.nnz)
        self.dmax = int(info.dmax)
        self.edges_read = int(info.edges_read)
        self.has_zero_id = bool(info.has_zero_id)
        self.multigraph = bool(info.multigraph)
        self._rp = info.rowptr
        self._col = info.col

    @property
    def rowptr(self):
        return np.ctypeslib.as_array(self._rp, shape=(self.V + 1,)).copy()

    @property
    def col(self):
        if self.nnz == 0:
            return np.zeros(0, np.uint32)
        return np.ctypeslib.as_array(self._col, shape=(self.nnz,)).copy()

    @property
    def m(self):
        return self.nnz // 2

    def vgraph(self, part_size=64):
        """Degree-prefixed layout of virtual_graph.cpp:93-183."""
        P = ctypes.c_uint64(0)
        _check(lib().kc_hg_vgraph(self._h, part_size, None, None, None, ctypes.byref(P)))
        npn = np.zeros(max(self.V, 1), np.uint32)
        el = np.zeros(max(self.nnz + self.V, 1), np.uint32)
        parts = np.zeros(max(2 * P.value, 2), np.uint32)
        _check(lib().kc_hg_vgraph(self._h, part_size, _ptr(npn), _ptr(el), _ptr(parts),
                                  ctypes.byref(P)))
        return npn[:self.V], el[:self.nnz + self.V], parts[:2 * P.value].reshape(-1, 2)

    def write_edgelist(self, path):
        _check(lib().kc_hg_write_edgelist(self._h, os.fsencode(path)))

    def close(self):
        if self._h:
            lib().kc_hg_free(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)

    @staticmethod
    def _make(fn, *args):
        h = _vp()
        _check(fn(*args, ctypes.byref(h)))
        return HostGraph(h)

    @classmethod
    def read_coo(cls, path):
        return cls._make(lib().kc_hg_read_coo, os.fsencode(path))

    @classmethod
    def read_dedup(cls, path):
        return cls._make(lib().kc_hg_read_dedup, os.fsencode(path))

    @classmethod
    def load(cls, path):
        """A graph written by save() (binary CSR)."""
        return cls._make(lib().kc_hg_load, os.fsencode(path))

    def save(self, path):
        _check(lib().kc_hg_save(self._h, os.fsencode(path)))

    @classmethod
    def read_cached(cls, path, reader, cache_dir):
        """Read `path` with reader "coo" or "dedup" through the content-keyed
        binary cache in `cache_dir`; returns (graph, hit)."""
        r = {"coo": KC_READER_COO, "dedup": KC_READER_DEDUP}[reader]
        h = _vp()
        hit = ctypes.c_int(0)
        _check(lib().kc_hg_read_cached(os.fsencode(path), r, os.fsencode(cache_dir),
                                       ctypes.byref(h), ctypes.byref(hit)))
        return cls(h), bool(hit.value)

    @classmethod
    def from_pairs(cls, V, src, dst):
        s = np.ascontiguousarray(src, np.uint32)
        d = np.ascontiguousarray(dst, np.uint32)
        return cls._make(lib().kc_hg_from_pairs, V, s.size, _ptr(s), _ptr(d))

    @classmethod
    def from_csr(cls, rowptr, col):
        rp = np.ascontiguousarray(rowptr, np.uint64)
        cl = np.ascontiguousarray(col, np.uint32)
        return cls._make(lib().kc_hg_from_csr, len(rp) - 1, _ptr(rp), _ptr(cl))

    @classmethod
    def er(cls, n, m, seed=0):
        return cls._make(lib().kc_hg_gen_er, n, m, seed)

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
        m_gen = int(round(edge_factor * (1 << scale)))
        return cls._make(lib().kc_hg_gen_rmat, scale, m_gen, a, b, c, seed)

    @classmethod
    def chunglu(cls, n, beta=2.1, mean_deg=20.0, cap=4e6, seed=3):
        return cls._make(lib().kc_hg_gen_chunglu, n, beta, mean_deg, cap, seed)


class HostBatch:
    """Many independent graphs in the concatenated layout (kc_host_batch)."""

    def __init__(self, handle):
        self._h = handle
        _register(self)
        info = BatchInfo()
        _check(lib().kc_hb_info(self._h, ctypes.byref(info)))
        self.G = int(info.G)
        self.vertices = int(info.vertices)
        self.entries = int(info.entries)
        self._info = info

    def _arr(self, ptr, n):
        if n == 0 or not ptr:
            return None
        return np.ctypeslib.as_array(ptr, shape=(n,))

    @property
    def vtx_off(self):
        return self._arr(self._info.vtx_off, self.G + 1).copy()

    @property
    def edge_off(self):
        return self._arr(self._info.edge_off, self.G + 1).copy()

    @property
    def scales(self):
        a = self._arr(self._info.scale, self.G)
        return None if a is None else a.copy()

    @property
    def m_gen(self):
        a = self._arr(self._info.m_gen, self.G)
        return None if a is None else a.copy()

    def arrays(self):
        """(vtx_off, edge_off, rowptr, col) views into the batch's memory."""
        return (self._arr(self._info.vtx_off, self.G + 1),
                self._arr(self._info.edge_off, self.G + 1),
                self._arr(self._info.rowptr, self.vertices + self.G),
                self._arr(self._info.col, self.entries))

    def graph(self, g):
        h = _vp()
        _check(lib().kc_hb_graph(self._h, g, ctypes.byref(h)))
        return HostGraph(h)

    def close(self):
        if self._h:
            lib().kc_hb_free(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)

    @classmethod
    def rmat(cls, G, scale_min=8, scale_max=18, ef_min=4.0, ef_max=32.0, a=0.57, b=0.19,
             c=0.19, seed=1):
        h = _vp()
        _check(lib().kc_hb_gen_rmat(G, scale_min, scale_max, ef_min, ef_max, a, b, c, seed,
                                    ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_graphs(cls, graphs):
        arr = (_vp * max(len(graphs), 1))(*[g._h for g in graphs])
        h = _vp()
        _check(lib().kc_hb_from_graphs(len(graphs), arr, ctypes.byref(h)))
        return cls(h)


# ---------------------------------------------------------------- device
class Context:
    """kc_ctx: a device and its stream.  Graphs and batches made on it hold
    a reference, so closing the context while they live is safe."""

    def __init__(self, device=0):
        self._h = _vp()
        _check(lib().kc_ctx_create(device, ctypes.byref(self._h)))
        _register(self)
        cus = ctypes.c_uint32(0)
        hbm = ctypes.c_uint64(0)
        _check(lib().kc_ctx_info(self._h, ctypes.byref(cus), ctypes.byref(hbm)))
        self.num_cus = cus.value
        self.hbm_bytes = hbm.value

    def close(self):
        if self._h:
            lib().kc_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)

    def upload(self, graph, rowptr_bytes=4):
        h = _vp()
        _check(lib().kc_graph_upload_hg(self._h, graph._h, rowptr_bytes, ctypes.byref(h)))
        return DeviceGraph(self, h, graph.V)

    def probe_copy(self, nbytes=1 << 30, reps=5):
        """Mean ms of a streaming copy of nbytes (reads and writes nbytes)."""
        ms = ctypes.c_double(0)
        _check(lib().kc_probe_copy(self._h, nbytes, reps, ctypes.byref(ms)))
        return ms.value

    def gen_rmat(self, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, rowptr_bytes=4):
        """HostGraph.rmat's graph, built in HBM."""
        m_gen = int(round(edge_factor * (1 << scale)))
        h = _vp()
        _check(lib().kc_graph_gen_rmat(self._h, scale, m_gen, a, b, c, seed, rowptr_bytes,
                                       ctypes.byref(h)))
        return DeviceGraph(self, h, 1 << scale)

    def gen_chunglu(self, n, beta=2.1, mean_deg=20.0, cap=4e6, seed=3, rowptr_bytes=4):
        """HostGraph.chunglu's graph, built in HBM."""
        h = _vp()
        _check(lib().kc_graph_gen_chunglu(self._h, n, beta, mean_deg, cap, seed, rowptr_bytes,
                                          ctypes.byref(h)))
        return DeviceGraph(self, h, n)

    def read(self, path, reader="coo", rowptr_bytes=4):
        """An edge file straight into a resident graph (kc_graph_read): the
        HostGraph.read_coo / read_dedup graph, parsed on host threads and
        built on the device.  Returns (DeviceGraph, stats dict)."""
        r = {"coo": 0, "dedup": 1}[reader] if isinstance(reader, str) else int(reader)
        h = _vp()
        st = ReadStats()
        _check(lib().kc_graph_read(self._h, os.fsencode(path), r, rowptr_bytes,
                                   ctypes.byref(h), ctypes.byref(st)))
        return DeviceGraph(self, h, int(st.V)), st.as_dict()

    def gen_rmat_batch(self, G, scale_min=8, scale_max=18, ef_min=4.0, ef_max=32.0, a=0.57,
                       b=0.19, c=0.19, seed=1):
        """HostBatch.rmat's batch, built in HBM."""
        h = _vp()
        _check(lib().kc_batch_gen_rmat(self._h, G, scale_min, scale_max, ef_min, ef_max, a, b,
                                       c, seed, ctypes.byref(h)))
        return DeviceBatch(self, handle=h)

    def upload_csr(self, rowptr, col, rowptr_bytes=4):
        rp = np.ascontiguousarray(rowptr, np.uint64)
        cl = np.ascontiguousarray(col, np.uint32)
        h = _vp()
        _check(lib().kc_graph_upload(self._h, len(rp) - 1, _ptr(rp), _ptr(cl), rowptr_bytes,
                                     ctypes.byref(h)))
        return DeviceGraph(self, h, len(rp) - 1)


class DegreeStats(ctypes.Structure):
    _fields_ = [("V", ctypes.c_uint32), ("dmax", ctypes.c_uint32), ("nnz", ctypes.c_uint64),
                ("mean", ctypes.c_double), ("stddev", ctypes.c_double),
                ("reserved", ctypes.c_uint64 * 2)]


class DeviceGraph:
    def degree_stats(self):
        """V, max degree, nnz, mean and population standard deviation of the
        degrees (what peelone_dyn prints: main.cu:10-63)."""
        ds = DegreeStats()
        _check(lib().kc_graph_degree_stats(self._h, ctypes.byref(ds)))
        return {k: getattr(ds, k) for k, _ in ds._fields_ if k != "reserved"}

    def kcore_subgraph(self, k):
        """The k-core (core >= k, from the last decomposition) as a new device
        graph, and the original id of each of its vertices."""
        ids = np.zeros(max(self.V, 1), np.uint32)
        n = ctypes.c_uint32(0)
        h = _vp()
        _check(lib().kc_kcore_subgraph(self._h, k, ctypes.byref(h), _ptr(ids), ctypes.byref(n)))
        return DeviceGraph(self.ctx, h, n.value), ids[:n.value]

    def autotune(self, method="peelone_dyn", reps=1):
        """Time the launch candidates of `method` on this graph and keep the
        fastest as the graph's default; returns (opts dict, ms)."""
        m = METHODS[method] if isinstance(method, str) else int(method)
        o = Opts()
        ms = ctypes.c_double(0)
        _check(lib().kc_autotune(self._h, m, reps, ctypes.byref(o), ctypes.byref(ms)))
        return {"block": o.block, "hub_split": o.hub_split, "local_cap": o.local_cap,
                "agg_div": o.agg_div, "compact_div": o.compact_div,
                "grid": o.grid, "stage_grid": o.stage_grid}, ms.value

    def __init__(self, ctx, handle, V):
        self.ctx = ctx
        self._h = handle
        self.V = V
        _register(self)

    def download(self):
        h = _vp()
        _check(lib().kc_graph_download(self._h, ctypes.byref(h)))
        return HostGraph(h)

    @property
    def nnz(self):
        n = ctypes.c_uint64(0)
        _check(lib().kc_graph_dims(self._h, None, ctypes.byref(n)))
        return n.value

    def update_prepare(self):
        """Build the maintenance state now rather than in the first update."""
        _check(lib().kc_update_prepare(self._h))

    def update(self, op, u, v, copy=True):
        """Insert (op "insert") or delete ("delete") the unordered pairs
        (u[i], v[i]) and return (maintained cores, stats dict)."""
        o = {"insert": 0, "delete": 1}[op] if isinstance(op, str) else int(op)
        uu = np.ascontiguousarray(u, np.uint32)
        vv = np.ascontiguousarray(v, np.uint32)
        if uu.size != vv.size:
            raise ValueError("u and v differ in length")
        core = np.zeros(max(self.V, 1), np.uint32) if copy else None
        st = UpdateStats()
        _check(lib().kc_update(self._h, o, _ptr(uu), _ptr(vv), uu.size, _ptr(core),
                               ctypes.byref(st)))
        return (core[:self.V] if copy else None), st.as_dict()

    @property
    def device_bytes(self):
        return int(lib().kc_graph_device_bytes(self._h))

    def decompose(self, method="peelone_dyn", copy=True, **opts):
        m = METHODS[method] if isinstance(method, str) else int(method)
        o = make_opts(**opts)
        st = Stats()
        core = np.zeros(max(self.V, 1), np.uint32) if copy else None
        _check(lib().kc_decompose(self._h, m, ctypes.byref(o), _ptr(core), ctypes.byref(st)))
        return (core[:self.V] if copy else None), st.as_dict()

    def close(self):
        if self._h:
            lib().kc_graph_free(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)


class DeviceBatch:
    """A batch resident on the device (kc_batch)."""

    def __init__(self, ctx, hb=None, arrays=None, handle=None):
        self.ctx = ctx
        h = _vp()
        if handle is not None:
            h = handle
            G = ctypes.c_uint32(0)
            nv = ctypes.c_uint64(0)
            _check(lib().kc_batch_dims(h, ctypes.byref(G), ctypes.byref(nv), None))
            self.G, self.vertices = G.value, nv.value
        elif hb is not None:
            _check(lib().kc_batch_upload_hb(ctx._h, hb._h, ctypes.byref(h)))
            self.G, self.vertices = hb.G, hb.vertices
        else:
            vo, eo, rp, cl = [np.ascontiguousarray(x, t) for x, t in
                              zip(arrays, (np.uint64, np.uint64, np.uint32, np.uint32))]
            _check(lib().kc_batch_upload(ctx._h, len(vo) - 1, _ptr(vo), _ptr(eo), _ptr(rp),
                                         _ptr(cl), ctypes.byref(h)))
            self.G, self.vertices = len(vo) - 1, int(vo[-1])
        self._h = h
        _register(self)

    @property
    def device_bytes(self):
        return int(lib().kc_batch_device_bytes(self._h))

    def run(self, block=0, per_cu=0, copy=True, lds_vertices=0, split=0):
        o = BatchOpts()
        o.block, o.per_cu, o.lds_vertices, o.split = block, per_cu, lds_vertices, split
        st = BatchStats()
        core = np.zeros(max(self.vertices, 1), np.uint32) if copy else None
        km = np.zeros(max(self.G, 1), np.uint32) if copy else None
        _check(lib().kc_batch_run(self._h, ctypes.byref(o), _ptr(core), _ptr(km),
                                  ctypes.byref(st)))
        if not copy:
            return None, None, st.as_dict()
        return core[:self.vertices], km[:self.G], st.as_dict()

    def run_plain(self, method="peelone_dyn", copy=True, **opts):
        m = METHODS[method] if isinstance(method, str) else int(method)
        o = make_opts(**opts)
        st = BatchStats()
        core = np.zeros(max(self.vertices, 1), np.uint32) if copy else None
        km = np.zeros(max(self.G, 1), np.uint32) if copy else None
        _check(lib().kc_batch_run_plain(self._h, m, ctypes.byref(o), _ptr(core), _ptr(km),
                                        ctypes.byref(st)))
        if not copy:
            return None, None, st.as_dict()
        return core[:self.vertices], km[:self.G], st.as_dict()

    def download(self):
        h = _vp()
        _check(lib().kc_batch_download(self._h, ctypes.byref(h)))
        return HostBatch(h)

    def autotune(self, reps=1):
        o = BatchOpts()
        ms = ctypes.c_double(0)
        _check(lib().kc_batch_autotune(self._h, reps, ctypes.byref(o), ctypes.byref(ms)))
        return {"block": o.block, "per_cu": o.per_cu, "lds_vertices": o.lds_vertices,
                "split": o.split}, ms.value

    def close(self):
        if self._h:
            lib().kc_batch_free(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)


def core_checksum(core):
    """kc_core_checksum: sum over v of splitmix64(v << 32 | core[v]) mod 2^64."""
    c = np.ascontiguousarray(core, np.uint32)
    return int(lib().kc_core_checksum(_ptr(c), c.size))


class Server:
    """kc_server: the request loop (admission by HBM budget, fused small
    decompositions, resident graphs updated in submission order)."""

    KIND = {"rmat": 0, "update": 1, "decompose": 2}

    def __init__(self, device=0, slots=4, max_batch=256, fuse_max_scale=16, hbm_budget=0):
        o = ServerOpts()
        o.slots, o.max_batch, o.fuse_max_scale, o.hbm_budget = (slots, max_batch,
                                                                fuse_max_scale, hbm_budget)
        self._h = _vp()
        _check(lib().kc_server_create(device, ctypes.byref(o), ctypes.byref(self._h)))
        _register(self)
        self._keep = {}

    def gen_rmat_graph(self, scale, edge_factor=16.0, seed=1):
        gid = ctypes.c_uint32(0)
        _check(lib().kc_server_gen_rmat_graph(self._h, scale, edge_factor, seed,
                                              ctypes.byref(gid)))
        return gid.value

    def download_graph(self, gid):
        h = _vp()
        _check(lib().kc_server_graph_download(self._h, gid, ctypes.byref(h)))
        return HostGraph(h)

    def _submit(self, r, keep):
        rid = ctypes.c_uint64(0)
        _check(lib().kc_server_submit(self._h, ctypes.byref(r), ctypes.byref(rid)))
        if keep is not None:
            self._keep[rid.value] = keep
        return rid.value

    def submit_rmat(self, scale, edge_factor=16.0, seed=1, tag=0, core_out=None):
        r = Request()
        r.kind, r.scale, r.edge_factor, r.seed, r.tag = 0, scale, edge_factor, seed, tag
        if core_out is not None:
            r.core_out = core_out.ctypes.data
        return self._submit(r, core_out)

    def submit_update(self, gid, op, u, v, tag=0, core_out=None):
        uu = np.ascontiguousarray(u, np.uint32)
        vv = np.ascontiguousarray(v, np.uint32)
        r = Request()
        r.kind, r.graph, r.op, r.tag = 1, gid, {"insert": 0, "delete": 1}[op], tag
        r.n, r.u, r.v = uu.size, uu.ctypes.data if uu.size else None, \
            vv.ctypes.data if vv.size else None
        if core_out is not None:
            r.core_out = core_out.ctypes.data
        return self._submit(r, core_out)

    def submit_decompose(self, gid, tag=0, core_out=None):
        r = Request()
        r.kind, r.graph, r.tag = 2, gid, tag
        if core_out is not None:
            r.core_out = core_out.ctypes.data
        return self._submit(r, core_out)

    def poll(self, max_n=256, wait_ms=-1):
        buf = (Completion * max_n)()
        n = ctypes.c_uint32(0)
        _check(lib().kc_server_poll(self._h, buf, max_n, ctypes.byref(n), wait_ms))
        out = []
        for i in range(n.value):
            d = buf[i].as_dict()
            self._keep.pop(d["id"], None)
            out.append(d)
        return out

    def drain(self):
        _check(lib().kc_server_drain(self._h))

    def now_ms(self):
        return float(lib().kc_server_now_ms(self._h))

    def stats(self):
        st = ServerStats()
        _check(lib().kc_server_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if self._h:
            lib().kc_server_destroy(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)


def vgraph_decompose(node_pointer, edge_list, method, **opts):
    """Device half of the virtual-graph programs on their native layout."""
    npn = np.ascontiguousarray(node_pointer, np.uint32)
    el = np.ascontiguousarray(edge_list, np.uint32)
    V = npn.size
    m = METHODS[method] if isinstance(method, str) else int(method)
    core = np.zeros(max(V, 1), np.uint32)
    st = Stats()
    o = make_opts(**opts)
    _check(lib().kc_vgraph_decompose(V, el.size - V, _ptr(npn), _ptr(el), m, ctypes.byref(o),
                                     _ptr(core), ctypes.byref(st)))
    return core[:V], st.as_dict()


def ref_peel(graph, which="peelone_dyn", want_cores=True):
    """Drop-in for ``int peel_dyn(Graph&)`` / ``int peelone_dyn(Graph&)``:
    returns (ms, kmax, cores) for a dedup-reader graph."""
    rp = np.ascontiguousarray(graph.rowptr, np.uint32)
    nb = np.ascontiguousarray(graph.col, np.uint32)
    deg = np.diff(rp).astype(np.uint32)
    rg = RefGraph()
    rg.V = graph.V
    rg.E = int(rp[-1])
    rg.neighbors = nb.ctypes.data_as(_u32p) if nb.size else None
    rg.neighbors_offset = rp.ctypes.data_as(_u32p)
    rg.degrees = deg.ctypes.data_as(_u32p) if deg.size else None
    rg.dmax = int(deg.max()) if deg.size else 0
    core = np.zeros(max(graph.V, 1), np.uint32) if want_cores else None
    fn = lib().kc_peel_dyn if which == "peel_dyn" else lib().kc_peelone_dyn
    ms = fn(ctypes.byref(rg), _ptr(core))
    if ms < 0:
        raise KCoreError(3, lib().kc_last_error().decode(errors="replace"))
    return ms, int(rg.kmax), (core[:graph.V] if want_cores else None)


def _trace_sig():
    L = lib()
    if not getattr(L.kc_graph_trace, "_kc_sig", False):
        L.kc_graph_trace.argtypes = [_vp, _vp, ctypes.c_uint32, _u64p]
        L.kc_graph_trace.restype = ctypes.c_uint32
        L.kc_graph_trace._kc_sig = True
    return L


KC_OPT_TRACE = 1
KC_OPT_NO_DEGSORT = 2
KC_OPT_COUNT = 4
KC_READER_COO = 0
KC_READER_DEDUP = 1


def device_trace(dgraph, max_phases=1 << 16):
    """Per-phase trace of the last decompose(..., flags=KC_OPT_TRACE):
    a dict of numpy arrays with times in milliseconds."""
    L = _trace_sig()
    buf = np.zeros((max_phases, 16), np.uint64)
    tpm = ctypes.c_uint64(0)
    n = L.kc_graph_trace(dgraph._h, buf.ctypes.data_as(_vp), max_phases, ctypes.byref(tpm))
    t = buf[:n].astype(np.float64)
    f = float(tpm.value or 1)
    t0 = t[0, 0] if n else 0.0
    return {
        "start_ms": (t[:, 0] - t0) / f, "wg0_done_ms": (t[:, 1] - t0) / f,
        "exit_ms": (t[:, 2] - t0) / f, "last_done_ms": (t[:, 3] - t0) / f,
        "level": buf[:n, 4].copy(), "claims": buf[:n, 5].copy(),
        "shared": buf[:n, 6].copy(), "kind": buf[:n, 7].copy(),
        "max_rounds": buf[:n, 8].copy(), "edges": buf[:n, 9].copy(),
        "wg0_rounds": buf[:n, 10].copy(), "max_wg_edges": buf[:n, 11].copy(),
        "scan_done_ms": (t[:, 12] - t0) / f, "own_done_ms": (t[:, 13] - t0) / f,
        "last_start_ms": (t[:, 14] - t0) / f,
    }


def device_histogram(dgraph, nwords):
    """HistoCore's final histogram (reference layout, nnz+V words)."""
    L = lib()
    if not getattr(L.kc_graph_histogram, "_kc_sig", False):
        L.kc_graph_histogram.argtypes = [_vp, _vp, ctypes.c_uint64]
        L.kc_graph_histogram.restype = ctypes.c_int
        L.kc_graph_histogram._kc_sig = True
    out = np.zeros(max(nwords, 1), np.uint32)
    _check(L.kc_graph_histogram(dgraph._h, out.ctypes.data_as(_vp), nwords))
    return out[:nwords]
