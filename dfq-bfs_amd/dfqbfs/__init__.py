"""ctypes view of the MI355X DFQ-BFS C ABI (include/dfq_graph.h, dfq_bfs.h,
dfq_batch.h, dfq_serve.h).

The product is the native library ``lib/libdfqbfs.so`` (C++ host code and
HIP kernels).  This module only binds its entry points for the tests and the
benchmark; it does no BFS arithmetic of its own and has no fallback: if the
library is missing, importing this module raises.

Names follow the reference host driver
(design/shared-q/atomic-sa/our-sep/upper-bound/bfs.cu:181-317): a graph is
loaded or generated on the host (``Csr``), made resident on the GPU
(``DeviceGraph``), and searched from a source; the result is the per-vertex
level array (``UNREACHED`` for vertices the search does not reach) and,
optionally, a BFS-tree parent array.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libdfqbfs.so")
CLI_PATH = os.path.join(PKG_ROOT, "bin", "bfs")

UNREACHED = 0xFFFFFFFF

OK = 0
ERR_INVALID_ARG = -1
ERR_IO = -2
ERR_FORMAT = -3
ERR_RANGE = -4
ERR_NOMEM = -5
ERR_HIP = -6
ERR_NO_DEVICE = -7
ERR_COOPERATIVE = -8
ERR_CAPACITY = -9
ERR_BUSY = -10
ERR_STATE = -11
ERR_PENDING = -12
ERR_TIMEOUT = -13

CSR_SYMMETRIZE = 0x1
CSR_DEDUP = 0x2
CSR_DROP_SELF = 0x4
CSR_PERMUTE = 0x8
CSR_UNDIRECTED = CSR_SYMMETRIZE | CSR_DEDUP | CSR_DROP_SELF | CSR_PERMUTE

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run `make -C {PKG_ROOT}` (or __graft_entry__.build())")

lib = C.CDLL(LIB_PATH)

LAUNCH_PLAIN, LAUNCH_COOPERATIVE, LAUNCH_PLAIN_ABANDON = 0, 1, 2


# Every device object still open when the interpreter exits is closed by an
# exit handler, newest first, while the library and its HIP runtime are
# whole: left to __del__ during interpreter teardown, the frees could run
# after a profiler attached to the process has shut down.
_open_objects: "list[weakref.ref]" = []


def _track(obj):
    _open_objects.append(weakref.ref(obj))
    if len(_open_objects) > 256:
        _open_objects[:] = [r for r in _open_objects if r() is not None]


@atexit.register
def close_all():
    """Closes every open DeviceGraph, DeviceBuffer, Batch and Server."""
    while _open_objects:
        obj = _open_objects.pop()()
        if obj is not None:
            obj.close()


class DfqError(RuntimeError):
    def __init__(self, code: int, what: str):
        self.code = code
        super().__init__(f"{what}: {lib.dfq_strerror(code).decode()} ({code})")


class _Csr(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint64),
                ("row_offsets", C.POINTER(C.c_int64)), ("col_idx", C.POINTER(C.c_uint32))]


class Config(C.Structure):
    _fields_ = [("block_size", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("small_max", C.c_uint32), ("medium_max", C.c_uint32),
                ("large_max", C.c_uint32), ("direction_opt", C.c_uint32),
                ("bu_enter_div", C.c_uint32), ("bu_exit_div", C.c_uint32),
                ("launch_mode", C.c_uint32), ("multi_bu_div", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("depth", C.c_uint32), ("grid_blocks", C.c_uint32),
                ("block_size", C.c_uint32), ("bottom_up_levels", C.c_uint32),
                ("reached", C.c_uint64), ("edges", C.c_uint64)]


class DeviceInfo(C.Structure):
    _fields_ = [("device", C.c_int32), ("cu_count", C.c_int32), ("xcc_count", C.c_int32),
                ("cooperative", C.c_int32), ("lds_per_cu", C.c_uint64), ("hbm_bytes", C.c_uint64),
                ("l2_bytes", C.c_uint64), ("arch", C.c_char * 64)]


class ValidateReport(C.Structure):
    _fields_ = [("errors", C.c_uint64), ("bad_source", C.c_uint64), ("bad_edge_span", C.c_uint64),
                ("missed_vertex", C.c_uint64), ("bad_parent", C.c_uint64), ("reached", C.c_uint64),
                ("depth", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = C.c_void_p
_u32p = C.POINTER(C.c_uint32)
_i64p = C.POINTER(C.c_int64)
_csrp = C.POINTER(_Csr)


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("dfq_strerror", C.c_char_p, C.c_int)
_sig("dfq_csr_free", None, _csrp)
_sig("dfq_csr_from_edges", C.c_int, _u32p, _u32p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64,
     C.POINTER(_csrp))
_sig("dfq_csr_from_arrays", C.c_int, _i64p, _u32p, C.c_uint32, C.c_uint64, C.POINTER(_csrp))
_sig("dfq_load_snap", C.c_int, C.c_char_p, C.c_uint32, C.POINTER(_csrp))
_sig("dfq_save_snap", C.c_int, C.c_char_p, _csrp)
_sig("dfq_load_dat", C.c_int, C.c_char_p, C.POINTER(_csrp), _u32p)
_sig("dfq_save_dat", C.c_int, C.c_char_p, _csrp, C.c_uint32)
_sig("dfq_load_csr_bin", C.c_int, C.c_char_p, C.POINTER(_csrp))
_sig("dfq_save_csr_bin", C.c_int, C.c_char_p, _csrp)
_sig("dfq_gen_amazon_like", C.c_int, C.c_uint32, C.c_uint64, C.POINTER(_csrp))
_sig("dfq_gen_rmat", C.c_int, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
     C.c_uint64, C.c_uint32, C.POINTER(_csrp))
_sig("dfq_gen_grid2d", C.c_int, C.c_uint32, C.c_uint32, C.c_double, C.c_uint64, C.POINTER(_csrp))
_sig("dfq_pick_sources", C.c_int, _csrp, C.c_uint32, C.c_uint64, _u32p)
_sig("dfq_reached_work", C.c_int, _csrp, _u32p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
     _u32p)
_sig("dfq_validate_bfs", C.c_int, _csrp, C.c_uint32, _u32p, _u32p, C.POINTER(ValidateReport))
_sig("dfq_compare_levels", C.c_int, C.c_uint32, _u32p, _u32p, C.POINTER(C.c_uint64))
_sig("dfq_config_default", C.c_int, C.POINTER(Config))
_sig("dfq_get_device_info", C.c_int, C.POINTER(DeviceInfo))
_sig("dfq_graph_create", C.c_int, _i64p, _u32p, C.c_uint32, C.c_uint64, C.POINTER(_P))
_sig("dfq_graph_destroy", None, _P)
_sig("dfq_graph_set_config", C.c_int, _P, C.POINTER(Config))
_sig("dfq_graph_get_config", C.c_int, _P, C.POINTER(Config))
_sig("dfq_graph_device_arrays", C.c_int, _P, C.POINTER(_P), C.POINTER(_P))
_sig("dfq_bfs", C.c_int, _P, C.c_uint32, _u32p, _u32p, C.POINTER(Stats))
_sig("dfq_bfs_device", C.c_int, _P, C.c_uint32, _P, _P, C.POINTER(Stats))
_sig("dfq_bfs_device_many", C.c_int, _P, C.c_uint32, _u32p, _P, _P, C.POINTER(Stats))
_sig("dfq_bfs_device_seq", C.c_int, _P, C.c_uint32, _u32p, _P, C.c_uint64, _P, C.c_uint64,
     C.POINTER(Stats), C.POINTER(C.c_double))
_sig("dfq_bfs_profile", C.c_int, _P, C.c_uint32, C.POINTER(C.c_uint64), C.c_uint32,
     C.POINTER(Stats))
_sig("dfq_bfs_device_multi", C.c_int, _P, C.c_uint32, _u32p, _P, C.c_uint64, C.POINTER(Stats),
     C.POINTER(C.c_double))
_sig("dfq_bfs_multi_profile", C.c_int, _P, C.c_uint32, _u32p, _P, C.c_uint64, C.POINTER(C.c_uint64),
     C.c_uint32, C.POINTER(Stats))
MULTI_MAX_SOURCES = 64
_sig("dfq_graph_is_symmetric", C.c_int, _P)
_sig("dfq_device_alloc", C.c_int, C.c_size_t, C.POINTER(_P))
_sig("dfq_device_free", C.c_int, _P)
_sig("dfq_copy_to_host", C.c_int, _P, _P, C.c_size_t)
_sig("dfq_copy_to_device", C.c_int, _P, _P, C.c_size_t)
_sig("dfq_device_synchronize", C.c_int)
_sig("dfq_stream_copy", C.c_int, _P, _P, C.c_size_t, C.POINTER(C.c_double))


def check(rc: int, what: str):
    if rc != OK:
        raise DfqError(rc, what)


def _u32(a):
    return a.ctypes.data_as(_u32p)


def _i64(a):
    return a.ctypes.data_as(_i64p)


class Csr:
    """Host CSR owned by the native library (freed with the object)."""

    def __init__(self, ptr):
        self._ptr = ptr

    @property
    def n(self) -> int:
        return self._ptr.contents.n

    @property
    def m(self) -> int:
        return self._ptr.contents.m

    @property
    def row_offsets(self) -> np.ndarray:
        c = self._ptr.contents
        return np.ctypeslib.as_array(c.row_offsets, shape=(c.n + 1,))

    @property
    def col_idx(self) -> np.ndarray:
        c = self._ptr.contents
        if c.m == 0:
            return np.zeros(0, np.uint32)
        return np.ctypeslib.as_array(c.col_idx, shape=(c.m,))

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_offsets)

    def __del__(self):
        if getattr(self, "_ptr", None):
            lib.dfq_csr_free(self._ptr)
            self._ptr = None

    # -- constructors ------------------------------------------------------
    @staticmethod
    def _new(fn, what, *args) -> "Csr":
        out = _csrp()
        check(fn(*args, C.byref(out)), what)
        return Csr(out)

    @classmethod
    def from_edges(cls, src, dst, n: int, flags: int = 0, perm_seed: int = 0) -> "Csr":
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        if src.shape != dst.shape:
            raise ValueError("src and dst differ in length")
        return cls._new(lib.dfq_csr_from_edges, "dfq_csr_from_edges", _u32(src), _u32(dst),
                        src.size, n, flags, perm_seed)

    @classmethod
    def from_arrays(cls, row_offsets, col_idx) -> "Csr":
        ro = np.ascontiguousarray(row_offsets, np.int64)
        ci = np.ascontiguousarray(col_idx, np.uint32)
        return cls._new(lib.dfq_csr_from_arrays, "dfq_csr_from_arrays", _i64(ro), _u32(ci),
                        ro.size - 1, ci.size)

    @classmethod
    def rmat(cls, scale: int, edgefactor: int = 16, a=0.57, b=0.19, c=0.19, seed: int = 1,
             flags: int = CSR_UNDIRECTED) -> "Csr":
        return cls._new(lib.dfq_gen_rmat, "dfq_gen_rmat", scale, edgefactor, a, b, c, seed, flags)

    @classmethod
    def amazon_like(cls, n: int = 262144, seed: int = 0xD0F1) -> "Csr":
        return cls._new(lib.dfq_gen_amazon_like, "dfq_gen_amazon_like", n, seed)

    @classmethod
    def grid2d(cls, width: int, height: int, drop_p: float = 0.0, seed: int = 4096) -> "Csr":
        return cls._new(lib.dfq_gen_grid2d, "dfq_gen_grid2d", width, height, drop_p, seed)

    @classmethod
    def load_snap(cls, path: str, flags: int = 0) -> "Csr":
        return cls._new(lib.dfq_load_snap, "dfq_load_snap", path.encode(), flags)

    @classmethod
    def load_csr_bin(cls, path: str) -> "Csr":
        return cls._new(lib.dfq_load_csr_bin, "dfq_load_csr_bin", path.encode())

    @classmethod
    def load_dat(cls, path: str):
        out = _csrp()
        src = C.c_uint32(0)
        check(lib.dfq_load_dat(path.encode(), C.byref(out), C.byref(src)), "dfq_load_dat")
        return Csr(out), src.value

    # -- writers / queries -------------------------------------------------
    def save_snap(self, path: str):
        check(lib.dfq_save_snap(path.encode(), self._ptr), "dfq_save_snap")

    def save_dat(self, path: str, source: int = 0):
        check(lib.dfq_save_dat(path.encode(), self._ptr, source), "dfq_save_dat")

    def save_csr_bin(self, path: str):
        check(lib.dfq_save_csr_bin(path.encode(), self._ptr), "dfq_save_csr_bin")

    def pick_sources(self, count: int, seed: int) -> np.ndarray:
        out = np.zeros(max(count, 1), np.uint32)
        check(lib.dfq_pick_sources(self._ptr, count, seed, _u32(out)), "dfq_pick_sources")
        return out[:count]

    def reached_work(self, level: np.ndarray):
        """(directed entries of reached vertices, reached vertices, depth)."""
        level = np.ascontiguousarray(level, np.uint32)
        e, v, d = C.c_uint64(), C.c_uint64(), C.c_uint32()
        check(lib.dfq_reached_work(self._ptr, _u32(level), C.byref(e), C.byref(v), C.byref(d)),
              "dfq_reached_work")
        return e.value, v.value, d.value

    def validate(self, source: int, level: np.ndarray, parent: np.ndarray | None = None) -> dict:
        level = np.ascontiguousarray(level, np.uint32)
        pp = None
        if parent is not None:
            parent = np.ascontiguousarray(parent, np.uint32)
            pp = _u32(parent)
        rep = ValidateReport()
        check(lib.dfq_validate_bfs(self._ptr, source, _u32(level), pp, C.byref(rep)),
              "dfq_validate_bfs")
        return rep.as_dict()


def default_config() -> Config:
    c = Config()
    check(lib.dfq_config_default(C.byref(c)), "dfq_config_default")
    return c


def device_info() -> dict:
    info = DeviceInfo()
    check(lib.dfq_get_device_info(C.byref(info)), "dfq_get_device_info")
    d = {k: getattr(info, k) for k, _ in info._fields_}
    d["arch"] = info.arch.decode()
    return d


@dataclass
class BfsResult:
    level: np.ndarray
    parent: np.ndarray | None
    kernel_ms: float
    depth: int
    reached: int
    grid_blocks: int
    block_size: int
    bottom_up_levels: int = 0
    edges: int = 0        # CSR entries of reached vertices (E_r)


class DeviceGraph:
    """A CSR resident in GPU memory with its BFS working set."""

    def __init__(self, csr: Csr | None = None, *, row_offsets=None, col_idx=None):
        if csr is not None:
            ro, ci, n, m = csr._ptr.contents.row_offsets, csr._ptr.contents.col_idx, csr.n, csr.m
        else:
            self._ro = np.ascontiguousarray(row_offsets, np.int64)
            self._ci = np.ascontiguousarray(col_idx, np.uint32)
            ro, ci, n, m = _i64(self._ro), _u32(self._ci), self._ro.size - 1, self._ci.size
        self.n, self.m = n, m
        h = _P()
        check(lib.dfq_graph_create(ro, ci, n, m, C.byref(h)), "dfq_graph_create")
        self._h = h
        _track(self)

    def close(self):
        if getattr(self, "_h", None):
            lib.dfq_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def config(self) -> Config:
        c = Config()
        check(lib.dfq_graph_get_config(self._h, C.byref(c)), "dfq_graph_get_config")
        return c

    def set_config(self, cfg: Config | None = None, **kw):
        c = cfg if cfg is not None else self.config()
        for k, v in kw.items():
            setattr(c, k, v)
        check(lib.dfq_graph_set_config(self._h, C.byref(c)), "dfq_graph_set_config")

    def bfs(self, source: int, parents: bool = False) -> BfsResult:
        level = np.empty(self.n, np.uint32)
        parent = np.empty(self.n, np.uint32) if parents else None
        st = Stats()
        check(lib.dfq_bfs(self._h, source, _u32(level), _u32(parent) if parents else None,
                          C.byref(st)), "dfq_bfs")
        return BfsResult(level, parent, st.kernel_ms, st.depth, st.reached, st.grid_blocks,
                         st.block_size, st.bottom_up_levels, st.edges)

    @property
    def symmetric(self) -> bool:
        return bool(lib.dfq_graph_is_symmetric(self._h))

    PROFILE_FIELDS = ("frontier", "edges", "probes", "atomics", "claims", "items", "bottom_up",
                      "level_ns", "busy_max_ns", "busy_sum_ns",
                      "bin_small", "bin_medium", "bin_large", "bin_huge",
                      "idle_small", "idle_medium", "idle_large", "idle_huge",
                      "edges_max_wg", "bu_candidates")

    def profile(self, source: int, max_levels: int = 64):
        """Per-level counters: list of dicts (levels beyond max_levels dropped)."""
        rows = np.zeros((max_levels, len(self.PROFILE_FIELDS)), np.uint64)
        st = Stats()
        check(lib.dfq_bfs_profile(self._h, source, rows.ctypes.data_as(C.POINTER(C.c_uint64)),
                                  max_levels, C.byref(st)), "dfq_bfs_profile")
        depth = min(st.depth, max_levels)
        return [dict(zip(self.PROFILE_FIELDS, map(int, rows[i]))) for i in range(depth)], st

    def bfs_device_many(self, sources, d_level: int, d_parent: int | None = None):
        """Searches from every source back to back (outputs overwritten in
        turn); per-search Stats with event-bracketed kernel times."""
        so = np.ascontiguousarray(sources, np.uint32)
        st = (Stats * max(so.size, 1))()
        check(lib.dfq_bfs_device_many(self._h, so.size, _u32(so), d_level, d_parent, st),
              "dfq_bfs_device_many")
        return list(st[:so.size])

    def bfs_device_seq(self, sources, d_level: int, level_stride: int = 0,
                       d_parent: int | None = None, parent_stride: int = 0):
        """Searches from every source in one launch; returns (per-search
        Stats with device-clock times, event time of the launch in ms)."""
        so = np.ascontiguousarray(sources, np.uint32)
        st = (Stats * max(so.size, 1))()
        ms = C.c_double()
        check(lib.dfq_bfs_device_seq(self._h, so.size, _u32(so), d_level, level_stride, d_parent,
                                     parent_stride, st, C.byref(ms)), "dfq_bfs_device_seq")
        return list(st[:so.size]), ms.value

    def bfs_device_multi(self, sources, d_level: int, level_stride: int | None = None):
        """Up to 64 searches in one traversal (search i writes d_level + i *
        level_stride, default n); returns (per-search Stats, launch ms)."""
        so = np.ascontiguousarray(sources, np.uint32)
        st = (Stats * max(so.size, 1))()
        ms = C.c_double()
        check(lib.dfq_bfs_device_multi(self._h, so.size, _u32(so), d_level,
                                       self.n if level_stride is None else level_stride, st,
                                       C.byref(ms)), "dfq_bfs_device_multi")
        return list(st[:so.size]), ms.value

    MULTI_PROFILE_FIELDS = ("frontier", "edges", "bottom_up", "active", "items", "commit_ns",
                            "expand_ns", "next")

    def multi_profile(self, sources, d_level: int, max_levels: int = 64):
        """Per-level counters of one multi-source search (list of dicts)."""
        so = np.ascontiguousarray(sources, np.uint32)
        rows = np.zeros((max_levels, len(self.MULTI_PROFILE_FIELDS)), np.uint64)
        st = (Stats * max(so.size, 1))()
        check(lib.dfq_bfs_multi_profile(self._h, so.size, _u32(so), d_level, self.n,
                                        rows.ctypes.data_as(C.POINTER(C.c_uint64)), max_levels, st),
              "dfq_bfs_multi_profile")
        depth = max(s.depth for s in st[:so.size])
        return ([dict(zip(self.MULTI_PROFILE_FIELDS, map(int, rows[i]))) for i in range(min(depth, max_levels))],
                list(st[:so.size]))

    def bfs_device(self, source: int, d_level: int, d_parent: int | None = None) -> Stats:
        """BFS into caller-owned device buffers (raw device addresses)."""
        st = Stats()
        check(lib.dfq_bfs_device(self._h, source, C.c_void_p(d_level),
                                 C.c_void_p(d_parent) if d_parent else None, C.byref(st)),
              "dfq_bfs_device")
        return st


class DeviceBuffer:
    """Device allocation owned through the library's HIP runtime."""

    def __init__(self, nbytes: int):
        p = _P()
        check(lib.dfq_device_alloc(nbytes, C.byref(p)), "dfq_device_alloc")
        self.ptr = p.value
        self.nbytes = nbytes
        _track(self)

    def to_numpy(self, dtype=np.uint32, count=None, offset: int = 0) -> np.ndarray:
        """`count` elements from element `offset` on (all the rest by default)."""
        dt = np.dtype(dtype)
        avail = self.nbytes // dt.itemsize - offset
        count = avail if count is None else count
        if offset < 0 or count < 0 or count > avail:
            raise ValueError("range outside the buffer")
        out = np.empty(count, dt)
        if count:
            check(lib.dfq_copy_to_host(out.ctypes.data_as(_P), C.c_void_p(self.ptr + offset * dt.itemsize),
                                       out.nbytes), "dfq_copy_to_host")
        return out

    def from_numpy(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than buffer")
        check(lib.dfq_copy_to_device(C.c_void_p(self.ptr), a.ctypes.data_as(_P), a.nbytes),
              "dfq_copy_to_device")

    def free(self):
        if getattr(self, "ptr", None):
            lib.dfq_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    close = free

    def __del__(self):
        self.free()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()


def synchronize():
    check(lib.dfq_device_synchronize(), "dfq_device_synchronize")


def stream_copy(dst: "DeviceBuffer", src: "DeviceBuffer", nbytes: int) -> float:
    """Copy nbytes device-to-device with the library's streaming kernel; ms."""
    ms = C.c_double()
    check(lib.dfq_stream_copy(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), nbytes, C.byref(ms)),
          "dfq_stream_copy")
    return ms.value


# ---- batched searches (include/dfq_batch.h) ---------------------------------

TIER_WORKGROUP = 0
TIER_GRID = 1
WG_MAX_VERTICES = 131072


class QueryPlan(C.Structure):
    _fields_ = [("tier", C.c_uint32), ("small_max", C.c_uint32), ("medium_max", C.c_uint32),
                ("large_max", C.c_uint32), ("direction_opt", C.c_uint32), ("alpha", C.c_uint32),
                ("beta", C.c_uint32), ("reserved", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class BatchStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("queries", C.c_uint32), ("grid_queries", C.c_uint32),
                ("wg_queries", C.c_uint32), ("graph_replay", C.c_uint32), ("block_size", C.c_uint32),
                ("grid_blocks", C.c_uint32), ("team_xcds", C.c_uint32), ("reserved", C.c_uint32)]


class TuneEntry(C.Structure):
    _fields_ = [("bucket", C.c_uint32), ("queries", C.c_uint32), ("plan", QueryPlan),
                ("cost_us", C.c_double), ("sub_batch_ms", C.c_double)]


class TuneReport(C.Structure):
    _fields_ = [("entries", C.c_uint32), ("block_size", C.c_uint32), ("candidates", C.c_uint32),
                ("team_xcds", C.c_uint32), ("entry", TuneEntry * 32)]


_sig("dfq_batch_create", C.c_int, C.c_uint32, _u32p, C.POINTER(C.c_uint64), _i64p, _u32p,
     C.POINTER(_P))
_sig("dfq_batch_destroy", None, _P)
_sig("dfq_batch_set_queries", C.c_int, _P, C.c_uint32, _u32p, _u32p)
_sig("dfq_batch_set_plan", C.c_int, _P, C.c_uint32, C.POINTER(QueryPlan))
_sig("dfq_batch_get_plan", C.c_int, _P, C.c_uint32, C.POINTER(QueryPlan))
_sig("dfq_batch_set_block_size", C.c_int, _P, C.c_uint32)
_sig("dfq_batch_set_team_xcds", C.c_int, _P, C.c_uint32)
_sig("dfq_batch_run", C.c_int, _P, C.POINTER(BatchStats))
_sig("dfq_batch_levels", C.c_int, _P, C.c_uint32, _u32p)
_sig("dfq_batch_result", C.c_int, _P, C.c_uint32, _u32p, C.POINTER(C.c_uint64))
_sig("dfq_batch_autotune", C.c_int, _P, C.c_uint32, C.POINTER(TuneReport))


class Batch:
    """Many resident graphs; many (graph, source) queries per launch."""

    def __init__(self, graphs):
        self.graphs = list(graphs)
        n = np.array([g.n for g in self.graphs], np.uint32)
        m = np.array([g.m for g in self.graphs], np.uint64)
        ro = np.concatenate([g.row_offsets for g in self.graphs]).astype(np.int64)
        ci = (np.concatenate([g.col_idx for g in self.graphs]).astype(np.uint32)
              if m.sum() else np.zeros(1, np.uint32))
        self.n = n
        h = _P()
        check(lib.dfq_batch_create(len(self.graphs), _u32(n), m.ctypes.data_as(C.POINTER(C.c_uint64)),
                                   _i64(ro), _u32(ci), C.byref(h)), "dfq_batch_create")
        self._h = h
        _track(self)
        self.n_queries = 0

    def close(self):
        if getattr(self, "_h", None):
            lib.dfq_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_queries(self, graph_ids, sources):
        gi = np.ascontiguousarray(graph_ids, np.uint32)
        so = np.ascontiguousarray(sources, np.uint32)
        check(lib.dfq_batch_set_queries(self._h, gi.size, _u32(gi), _u32(so)), "dfq_batch_set_queries")
        self.q_graph, self.q_source, self.n_queries = gi.copy(), so.copy(), gi.size

    def plan(self, bucket: int) -> QueryPlan:
        p = QueryPlan()
        check(lib.dfq_batch_get_plan(self._h, bucket, C.byref(p)), "dfq_batch_get_plan")
        return p

    def set_plan(self, bucket: int, **kw):
        p = self.plan(bucket)
        for k, v in kw.items():
            setattr(p, k, v)
        check(lib.dfq_batch_set_plan(self._h, bucket, C.byref(p)), "dfq_batch_set_plan")

    def set_block_size(self, block: int):
        check(lib.dfq_batch_set_block_size(self._h, block), "dfq_batch_set_block_size")

    def set_team_xcds(self, xcds_per_team: int):
        """Grid-tier queries on teams of this many XCDs, several at once (0: the whole grid)."""
        check(lib.dfq_batch_set_team_xcds(self._h, xcds_per_team), "dfq_batch_set_team_xcds")

    def run(self) -> BatchStats:
        st = BatchStats()
        check(lib.dfq_batch_run(self._h, C.byref(st)), "dfq_batch_run")
        return st

    def levels(self, q: int) -> np.ndarray:
        out = np.empty(int(self.n[self.q_graph[q]]), np.uint32)
        check(lib.dfq_batch_levels(self._h, q, _u32(out)), "dfq_batch_levels")
        return out

    def result(self, q: int):
        d, r = C.c_uint32(), C.c_uint64()
        check(lib.dfq_batch_result(self._h, q, C.byref(d), C.byref(r)), "dfq_batch_result")
        return d.value, r.value

    def autotune(self, repeats: int = 1):
        rep = TuneReport()
        check(lib.dfq_batch_autotune(self._h, repeats, C.byref(rep)), "dfq_batch_autotune")
        return {"block_size": rep.block_size, "candidates": rep.candidates, "team_xcds": rep.team_xcds,
                "buckets": [dict(bucket=e.bucket, queries=e.queries, plan=e.plan.as_dict(),
                                 cost_us=e.cost_us, sub_batch_ms=e.sub_batch_ms)
                            for e in rep.entry[:rep.entries]]}


# ---------------------------------------------------------------------------
# Serving (include/dfq_serve.h)
# ---------------------------------------------------------------------------
SERVE_PARENTS = 0x1


class ServeConfig(C.Structure):
    _fields_ = [("block_size", C.c_uint32), ("xcds_per_team", C.c_uint32), ("ring_slots", C.c_uint32),
                ("groups", C.c_uint32), ("arena_bytes", C.c_uint64)]


class ServerInfo(C.Structure):
    _fields_ = [("running", C.c_uint32), ("teams", C.c_uint32), ("grid_blocks", C.c_uint32),
                ("block_size", C.c_uint32), ("xcds_per_team", C.c_uint32), ("ring_slots", C.c_uint32),
                ("in_flight", C.c_uint32), ("reserved", C.c_uint32), ("arena_bytes", C.c_uint64),
                ("arena_used", C.c_uint64), ("team_blocks", C.c_uint32 * 16)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("reserved", "team_blocks")}
        d["team_blocks"] = list(self.team_blocks[:self.teams])
        return d


class ServeResult(C.Structure):
    _fields_ = [("graph", C.c_uint32), ("source", C.c_uint32), ("depth", C.c_uint32),
                ("bottom_up_levels", C.c_uint32), ("reached", C.c_uint64), ("edges", C.c_uint64),
                ("team", C.c_uint32), ("reserved", C.c_uint32), ("service_us", C.c_double)]


class ServeRecord(C.Structure):
    _fields_ = [("arrival_us", C.c_double), ("admit_us", C.c_double), ("done_us", C.c_double),
                ("service_us", C.c_double), ("reached", C.c_uint64), ("edges", C.c_uint64),
                ("depth", C.c_uint32), ("team", C.c_uint32)]


_sig("dfq_serve_config_default", C.c_int, C.POINTER(ServeConfig))
_sig("dfq_serve_footprint", C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32)
_sig("dfq_serve_group_footprint", C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32)
_sig("dfq_server_submit_group", C.c_int, _P, C.c_uint32, C.c_uint32, _u32p, C.c_uint32,
     C.POINTER(C.c_uint64))
SERVE_GROUP_MAX = 64
_sig("dfq_server_create", C.c_int, C.c_uint32, _u32p, C.POINTER(C.c_uint64), _i64p, _u32p,
     C.POINTER(ServeConfig), C.POINTER(_P))
_sig("dfq_server_destroy", None, _P)
_sig("dfq_server_set_plan", C.c_int, _P, C.c_uint32, C.POINTER(QueryPlan))
_sig("dfq_server_get_plan", C.c_int, _P, C.c_uint32, C.POINTER(QueryPlan))
_sig("dfq_server_set_shape", C.c_int, _P, C.c_uint32, C.c_uint32)
_sig("dfq_server_get_info", C.c_int, _P, C.POINTER(ServerInfo))
_sig("dfq_server_start", C.c_int, _P)
_sig("dfq_server_stop", C.c_int, _P)
_sig("dfq_server_submit", C.c_int, _P, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64))
_sig("dfq_server_poll", C.c_int, _P, C.c_uint64, C.POINTER(ServeResult))
_sig("dfq_server_wait", C.c_int, _P, C.c_uint64, C.c_uint32, C.POINTER(ServeResult))
_sig("dfq_server_levels", C.c_int, _P, C.c_uint64, _u32p, _u32p)
_sig("dfq_server_release", C.c_int, _P, C.c_uint64)
_sig("dfq_server_replay", C.c_int, _P, C.c_uint32, _u32p, _u32p, C.POINTER(C.c_double), C.c_uint32,
     C.POINTER(_u32p), C.POINTER(ServeRecord))


def default_serve_config() -> ServeConfig:
    c = ServeConfig()
    check(lib.dfq_serve_config_default(C.byref(c)), "dfq_serve_config_default")
    return c


def serve_footprint(n: int, m: int, large_max: int = 256, xcds_per_team: int = 1, flags: int = 0) -> int:
    return int(lib.dfq_serve_footprint(n, m, large_max, xcds_per_team, flags))


class Server:
    """Resident graphs served by one persistent launch (continuous batching)."""

    def __init__(self, graphs, *, xcds_per_team: int | None = None, block_size: int | None = None,
                 ring_slots: int | None = None, arena_bytes: int | None = None, groups: bool = False):
        self.graphs = list(graphs)
        cfg = default_serve_config()
        cfg.groups = 1 if groups else 0
        if xcds_per_team is not None:
            cfg.xcds_per_team = xcds_per_team
        if block_size is not None:
            cfg.block_size = block_size
        if ring_slots is not None:
            cfg.ring_slots = ring_slots
        if arena_bytes is not None:
            cfg.arena_bytes = arena_bytes
        n = np.array([g.n for g in self.graphs], np.uint32)
        m = np.array([g.m for g in self.graphs], np.uint64)
        ro = np.concatenate([g.row_offsets for g in self.graphs]).astype(np.int64)
        ci = (np.concatenate([g.col_idx for g in self.graphs]).astype(np.uint32)
              if m.sum() else np.zeros(1, np.uint32))
        self.n = n
        h = _P()
        check(lib.dfq_server_create(len(self.graphs), _u32(n), m.ctypes.data_as(C.POINTER(C.c_uint64)),
                                    _i64(ro), _u32(ci), C.byref(cfg), C.byref(h)), "dfq_server_create")
        self._h = h
        _track(self)

    def close(self):
        if getattr(self, "_h", None):
            lib.dfq_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def info(self) -> dict:
        i = ServerInfo()
        check(lib.dfq_server_get_info(self._h, C.byref(i)), "dfq_server_get_info")
        return i.as_dict()

    def plan(self, graph: int) -> QueryPlan:
        p = QueryPlan()
        check(lib.dfq_server_get_plan(self._h, graph, C.byref(p)), "dfq_server_get_plan")
        return p

    def set_plan(self, graph: int, **kw):
        p = self.plan(graph)
        for k, v in kw.items():
            setattr(p, k, v)
        check(lib.dfq_server_set_plan(self._h, graph, C.byref(p)), "dfq_server_set_plan")

    def set_shape(self, xcds_per_team: int, block_size: int = 1024):
        check(lib.dfq_server_set_shape(self._h, xcds_per_team, block_size), "dfq_server_set_shape")

    def start(self):
        check(lib.dfq_server_start(self._h), "dfq_server_start")

    def stop(self):
        check(lib.dfq_server_stop(self._h), "dfq_server_stop")

    def submit(self, graph: int, source: int, flags: int = 0) -> int:
        """Ticket of the admitted query; raises DfqError (code ERR_BUSY) when refused for now."""
        t = C.c_uint64()
        check(lib.dfq_server_submit(self._h, graph, source, flags, C.byref(t)), "dfq_server_submit")
        return t.value

    def submit_group(self, graph: int, sources) -> list:
        """Group-mode servers: up to 64 queries of one graph as one group task;
        their tickets (DfqError with code ERR_BUSY when refused for now)."""
        so = np.ascontiguousarray(sources, np.uint32)
        t = (C.c_uint64 * max(so.size, 1))()
        check(lib.dfq_server_submit_group(self._h, graph, so.size, _u32(so), 0, t), "dfq_server_submit_group")
        return list(t[:so.size])

    def poll(self, ticket: int):
        r = ServeResult()
        rc = lib.dfq_server_poll(self._h, ticket, C.byref(r))
        if rc == ERR_PENDING:
            return None
        check(rc, "dfq_server_poll")
        return r

    def wait(self, ticket: int, timeout_ms: int = 60000) -> ServeResult:
        r = ServeResult()
        check(lib.dfq_server_wait(self._h, ticket, timeout_ms, C.byref(r)), "dfq_server_wait")
        return r

    def levels(self, ticket: int, graph: int, parents: bool = False):
        lv = np.empty(int(self.n[graph]), np.uint32)
        pa = np.empty(int(self.n[graph]), np.uint32) if parents else None
        check(lib.dfq_server_levels(self._h, ticket, _u32(lv), _u32(pa) if parents else None),
              "dfq_server_levels")
        return (lv, pa) if parents else lv

    def release(self, ticket: int):
        check(lib.dfq_server_release(self._h, ticket), "dfq_server_release")

    def replay(self, graph_ids, sources, arrival_us, flags: int = 0, keep_levels: bool = False):
        """Serves a trace; returns (records as a structured array, levels or None)."""
        gi = np.ascontiguousarray(graph_ids, np.uint32)
        so = np.ascontiguousarray(sources, np.uint32)
        ar = np.ascontiguousarray(arrival_us, np.float64)
        nq = gi.size
        recs = (ServeRecord * max(nq, 1))()
        outs = None
        out_ptrs = None
        if keep_levels:
            outs = [np.empty(int(self.n[g]), np.uint32) for g in gi]
            out_ptrs = (_u32p * max(nq, 1))(*[_u32(o) for o in outs])
        check(lib.dfq_server_replay(self._h, nq, _u32(gi), _u32(so), ar.ctypes.data_as(C.POINTER(C.c_double)),
                                    flags, out_ptrs, recs), "dfq_server_replay")
        dt = np.dtype([("arrival_us", "f8"), ("admit_us", "f8"), ("done_us", "f8"), ("service_us", "f8"),
                       ("reached", "u8"), ("edges", "u8"), ("depth", "u4"), ("team", "u4")])
        arr = np.frombuffer(bytearray(recs), dtype=dt, count=nq).copy() if nq else np.zeros(0, dt)
        return arr, outs
