"""dfq_bfs_device_noalloc (include/dfq_bfs.h): the BFS entry point that
allocates nothing, its scratch in one static device array per GPU.  Called
through the C ABI and compared with dfq_bfs_device (the existing entry point
it varies) on the same CSR: levels bit for bit, depth and reach equal,
parents certified; two calls in a row, the second on a smaller graph."""
import ctypes as C

import numpy as np
import pytest

from conftest import UNREACHED

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def noalloc(dfq):
    f = dfq.lib.dfq_bfs_device_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p,
                  C.POINTER(dfq.Stats)]
    return f


MAX_N = 1 << 18
MAX_M = 1 << 23


def run_both(dfq, csr, source, parents=True):
    """(levels, parents, stats) of the no-alloc call and of dfq_bfs_device."""
    f = noalloc(dfq)
    n, m = csr.n, csr.m
    with dfq.DeviceBuffer(8 * (n + 1)) as d_off, dfq.DeviceBuffer(4 * max(m, 1)) as d_col, \
            dfq.DeviceBuffer(4 * n) as d_lv, dfq.DeviceBuffer(4 * n) as d_pa:
        d_off.from_numpy(np.ascontiguousarray(csr.row_offsets, np.int64))
        if m:
            d_col.from_numpy(np.ascontiguousarray(csr.col_idx, np.uint32))
        st = dfq.Stats()
        rc = f(d_off.ptr, d_col.ptr, n, m, source, d_lv.ptr, d_pa.ptr if parents else None, C.byref(st))
        assert rc == dfq.OK, dfq.lib.dfq_strerror(rc)
        lv, pa = d_lv.to_numpy(), d_pa.to_numpy() if parents else None
    with dfq.DeviceGraph(csr) as g, dfq.DeviceBuffer(4 * n) as e_lv, dfq.DeviceBuffer(4 * n) as e_pa:
        est = g.bfs_device(source, e_lv.ptr, e_pa.ptr if parents else None)
        elv, epa = e_lv.to_numpy(), e_pa.to_numpy() if parents else None
    return (lv, pa, st), (elv, epa, est)


def check_same(csr, source, got, ref):
    lv, pa, st = got
    elv, epa, est = ref
    np.testing.assert_array_equal(lv, elv, err_msg=f"source {source}")
    assert (st.depth, st.reached) == (est.depth, est.reached)
    assert st.reached == int((lv != UNREACHED).sum())
    if pa is not None:
        assert csr.validate(source, lv, pa)["errors"] == 0
        assert csr.validate(source, elv, epa)["errors"] == 0


def test_noalloc_matches_dfq_bfs_device_two_calls(dfq, oracle):
    # first call: the largest graph the array holds room for (config-1 size),
    # second call: a smaller one, so stale state of the first would show
    big = dfq.Csr.amazon_like(MAX_N, seed=0xD0F1)
    small = dfq.Csr.rmat(14, 16, seed=41)
    assert big.n == MAX_N and big.m <= MAX_M
    for name, csr, s in (("amazon-like 262144", big, 0), ("rmat14", small, int(small.pick_sources(1, 3)[0]))):
        say("no-alloc vs dfq_bfs_device:", name, "source", s)
        got, ref = run_both(dfq, csr, s)
        check_same(csr, s, got, ref)
        np.testing.assert_array_equal(got[0], oracle.bfs(csr.row_offsets, csr.col_idx, s))
    # and again on the big graph, without parents, from another source
    s = int(big.pick_sources(1, 9)[0])
    say("no-alloc third call, source", s)
    got, ref = run_both(dfq, big, s, parents=False)
    check_same(big, s, got, ref)


def test_noalloc_hub_rows_and_limits(dfq, oracle):
    """Rows cut into work items (hubs past 256 entries) at the entry limit,
    then the refusals: sizes past the limits, a malformed CSR."""
    rng = np.random.default_rng(5)
    n = 200_000
    hubs = [0, 1, 2, 3]
    src, dst = [], []
    for h, d in zip(hubs, (150_000, 40_000, 257, 5_000)):
        t = rng.choice(n, size=d, replace=False)
        src += [h] * d
        dst += t.tolist()
    src += rng.integers(0, n, 600_000).tolist()
    dst += rng.integers(0, n, 600_000).tolist()
    csr = dfq.Csr.from_edges(src, dst, n, dfq.CSR_SYMMETRIZE | dfq.CSR_DEDUP | dfq.CSR_DROP_SELF)
    assert csr.m <= MAX_M
    for s in (0, 3, int(csr.pick_sources(1, 2)[0])):
        say("hubs, source", s)
        got, ref = run_both(dfq, csr, s)
        check_same(csr, s, got, ref)
        np.testing.assert_array_equal(got[0], oracle.bfs(csr.row_offsets, csr.col_idx, s))
    f = noalloc(dfq)
    st = dfq.Stats()
    with dfq.DeviceBuffer(8 * 16) as d_off, dfq.DeviceBuffer(64) as d_col, dfq.DeviceBuffer(64) as d_lv:
        assert f(d_off.ptr, d_col.ptr, MAX_N + 1, 1, 0, d_lv.ptr, None, C.byref(st)) == dfq.ERR_CAPACITY
        assert f(d_off.ptr, d_col.ptr, 4, MAX_M + 1, 0, d_lv.ptr, None, C.byref(st)) == dfq.ERR_CAPACITY
        assert f(d_off.ptr, d_col.ptr, 4, 4, 4, d_lv.ptr, None, C.byref(st)) == dfq.ERR_RANGE
        assert f(None, d_col.ptr, 4, 4, 0, d_lv.ptr, None, C.byref(st)) == dfq.ERR_INVALID_ARG
        # column index past n: refused by the device-side check, no search runs
        d_off.from_numpy(np.array([0, 1, 2, 3, 4], np.int64))
        d_col.from_numpy(np.array([1, 2, 3, 9], np.uint32))
        assert f(d_off.ptr, d_col.ptr, 4, 4, 0, d_lv.ptr, None, C.byref(st)) == dfq.ERR_RANGE
        # decreasing offsets: refused
        d_off.from_numpy(np.array([0, 3, 2, 3, 4], np.int64))
        d_col.from_numpy(np.array([1, 2, 3, 0], np.uint32))
        assert f(d_off.ptr, d_col.ptr, 4, 4, 0, d_lv.ptr, None, C.byref(st)) == dfq.ERR_RANGE
        # the same array, well formed: a 4-cycle
        d_off.from_numpy(np.array([0, 1, 2, 3, 4], np.int64))
        assert f(d_off.ptr, d_col.ptr, 4, 4, 0, d_lv.ptr, None, C.byref(st)) == dfq.OK
        np.testing.assert_array_equal(d_lv.to_numpy(count=4), [0, 1, 2, 3])
        assert st.depth == 4 and st.reached == 4
