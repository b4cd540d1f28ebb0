"""Shared fixtures: the native library binding, the oracle, small graphs.

The oracle (oracle/liboracle.so, a C restatement of the reference's serial
CPU BFS) is loaded here for use as the checker only.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dfq-bfs_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)
if ROOT not in sys.path:
    sys.path.insert(1, ROOT)  # bench.py: tests reuse its workload definitions
GOLDEN = os.path.join(ROOT, "tests", "golden")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
UNREACHED = 0xFFFFFFFF


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")
    config.addinivalue_line("markers", "slow: full-size inputs")


class Oracle:
    """ctypes view of oracle/liboracle.so."""

    def __init__(self, path=ORACLE_LIB):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        self.lib = C.CDLL(path)
        u32p = C.POINTER(C.c_uint32)
        self.lib.oracle_serial_bfs.restype = C.c_int
        self.lib.oracle_serial_bfs.argtypes = [C.c_uint32, C.POINTER(C.c_int64), u32p, C.c_uint32,
                                               u32p, u32p]
        self.lib.oracle_compare.restype = None
        self.lib.oracle_compare.argtypes = [C.c_uint32, u32p, u32p, u32p, C.POINTER(C.c_uint64),
                                            C.POINTER(C.c_uint64)]

    def bfs(self, row_offsets, col_idx, source, parents=False):
        ro = np.ascontiguousarray(row_offsets, np.int64)
        ci = np.ascontiguousarray(col_idx, np.uint32)
        n = ro.size - 1
        level = np.empty(n, np.uint32)
        parent = np.empty(n, np.uint32) if parents else None
        u32p = C.POINTER(C.c_uint32)
        rc = self.lib.oracle_serial_bfs(
            n, ro.ctypes.data_as(C.POINTER(C.c_int64)),
            ci.ctypes.data_as(u32p) if ci.size else C.cast(C.c_void_p(0), u32p), source,
            level.ctypes.data_as(u32p), parent.ctypes.data_as(u32p) if parents else None)
        assert rc == 0, rc
        return (level, parent) if parents else level

    def compare(self, cpu_level, gpu_level):
        u32p = C.POINTER(C.c_uint32)
        cpu = np.ascontiguousarray(cpu_level, np.uint32)
        gpu = np.ascontiguousarray(gpu_level, np.uint32)
        dist, touch, err = C.c_uint32(), C.c_uint64(), C.c_uint64()
        self.lib.oracle_compare(cpu.size, cpu.ctypes.data_as(u32p), gpu.ctypes.data_as(u32p),
                                C.byref(dist), C.byref(touch), C.byref(err))
        return dist.value, touch.value, err.value


@pytest.fixture(scope="session")
def oracle():
    return Oracle()


@pytest.fixture(scope="session")
def dfq():
    import dfqbfs
    return dfqbfs


def load_golden(name):
    """Loads tests/golden/<name>.npz (plain arrays, no pickle)."""
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_names():
    return sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))


def csr_from_edges_np(n, src, dst):
    """Plain numpy CSR (rows sorted, duplicates kept) for building test inputs."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    ro = np.zeros(n + 1, np.int64)
    np.add.at(ro, src + 1, 1)
    return np.cumsum(ro), dst.astype(np.uint32)


def gpu_available():
    try:
        import dfqbfs
        dfqbfs.device_info()
        return True
    except Exception:
        return False
