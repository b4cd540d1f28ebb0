"""dfq_bfs_device_multi (include/dfq_bfs.h): up to 64 searches of one graph
in one bit-parallel traversal.  Every search's levels equal the oracle's
(the serial CPU BFS restating the reference's host loop) bit for bit, its
reached count and depth match, whatever the mix of sources, repeats and
directions."""
import numpy as np
import pytest

from conftest import UNREACHED, load_golden

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def run_multi(dfq, g, srcs):
    n = g.n
    with dfq.DeviceBuffer(4 * n * len(srcs)) as buf:
        stats, ms = g.bfs_device_multi(srcs, buf.ptr)
        lv = [buf.to_numpy(count=n, offset=i * n) for i in range(len(srcs))]
    return lv, stats, ms


def check_multi(dfq, oracle, csr, srcs, direction_opt=1, expect=None):
    with dfq.DeviceGraph(csr) as g:
        g.set_config(direction_opt=direction_opt)
        lv, stats, ms = run_multi(dfq, g, srcs)
    say(f"  {len(srcs)} searches in {ms:.3f} ms, bottom-up levels {stats[0].bottom_up_levels}")
    for i, s in enumerate(srcs):
        e = expect[i] if expect is not None else oracle.bfs(csr.row_offsets, csr.col_idx, int(s))
        np.testing.assert_array_equal(lv[i], e, err_msg=f"search {i} source {s}")
        reached = e != UNREACHED
        assert stats[i].reached == int(reached.sum()), (i, stats[i].reached)
        assert stats[i].depth == int(e[reached].max()) + 1, (i, stats[i].depth)
    return stats


@pytest.mark.parametrize("count", [1, 2, 37, 64])
def test_multi_rmat_sources(dfq, oracle, count):
    csr = dfq.Csr.rmat(14, 16, seed=78)
    srcs = csr.pick_sources(count, seed=count)
    say("rmat14", count)
    check_multi(dfq, oracle, csr, srcs)


def test_multi_repeated_and_isolated_sources(dfq, oracle):
    """Repeated sources share a frontier vertex; a source of degree 0
    reaches only itself."""
    csr = dfq.Csr.rmat(12, 8, seed=5)
    deg = csr.degrees()
    iso = int(np.flatnonzero(deg == 0)[0])
    s = csr.pick_sources(20, seed=3)
    srcs = np.concatenate([s[:10], s[:5], [iso], s[10:], [iso, int(s[0])]]).astype(np.uint32)
    assert len(srcs) <= 64
    check_multi(dfq, oracle, csr, srcs)


@pytest.mark.parametrize("direction_opt", [0, 1])
def test_multi_directions_on_golden(dfq, oracle, direction_opt):
    """Golden graphs (scipy hop distances): top-down only, and with
    bottom-up levels where the graph is symmetric."""
    for name in ("rmat12_ef8", "bin_hubs_50000", "amazon_like_8192", "er_directed_3000"):
        gd = load_golden(name)
        csr = dfq.Csr.from_arrays(gd["row_offsets"], gd["col_idx"])
        say("golden", name, "direction_opt", direction_opt)
        check_multi(dfq, oracle, csr, gd["sources"], direction_opt, expect=gd["levels"])


def test_multi_hubs_past_every_bin(dfq, oracle):
    """Hubs whose rows go through the wave, workgroup and work-item bins of
    the top-down expansion, as sources and as neighbours."""
    gd = load_golden("bin_hubs_50000")
    csr = dfq.Csr.from_arrays(gd["row_offsets"], gd["col_idx"])
    deg = csr.degrees()
    hubs = np.argsort(deg)[-20:].astype(np.uint32)
    srcs = np.concatenate([hubs, csr.pick_sources(44, seed=9)]).astype(np.uint32)
    say("hubs", deg[hubs].tolist())
    check_multi(dfq, oracle, csr, srcs, direction_opt=0)
    check_multi(dfq, oracle, csr, srcs, direction_opt=1)


def test_multi_grid_depth_limit(dfq, oracle):
    """A 100 x 20 grid from a corner (118 levels) is covered; a 300-vertex
    path (300 levels) is past the 255-level planes and refused."""
    csr = dfq.Csr.grid2d(100, 20, 0.0)
    check_multi(dfq, oracle, csr, [0, 99, 1999, 1000, 0])
    n = 300
    path = dfq.Csr.from_edges(np.arange(n - 1), np.arange(1, n), n, dfq.CSR_SYMMETRIZE)
    with dfq.DeviceGraph(path) as g, dfq.DeviceBuffer(4 * n) as buf:
        with pytest.raises(dfq.DfqError) as e:
            g.bfs_device_multi([0], buf.ptr)
        assert e.value.code == dfq.ERR_CAPACITY
        # the same graph still searches correctly afterwards
        st = g.bfs_device(0, buf.ptr)
        assert st.depth == n


def test_multi_rmat22_64_sources(dfq, oracle):
    """The config-2 graph and the bench's 64 sources in one traversal: 8
    searches against the oracle, the rest against the certificate."""
    csr = dfq.Csr.rmat(22, 16, seed=22)
    srcs = csr.pick_sources(64, seed=2022)
    n = csr.n
    with dfq.DeviceGraph(csr) as g, dfq.DeviceBuffer(4 * n * 64) as buf:
        g.set_config(direction_opt=1)
        for rep in range(2):
            stats, ms = g.bfs_device_multi(srcs, buf.ptr)
            say(f"rmat22 64 searches: {ms:.3f} ms ({ms / 64 * 1e3:.1f} us per search), "
                f"bottom-up levels {stats[0].bottom_up_levels}")
        for i, s in enumerate(srcs):
            lv = buf.to_numpy(count=n, offset=i * n)
            if i % 8 == 0:
                np.testing.assert_array_equal(lv, oracle.bfs(csr.row_offsets, csr.col_idx, int(s)),
                                              err_msg=f"search {i}")
            else:
                assert csr.validate(int(s), lv)["errors"] == 0, i
            assert stats[i].reached == int((lv != UNREACHED).sum())


def test_multi_then_single_searches_share_no_state(dfq, oracle):
    """Multi-source and single-source launches alternate on one graph."""
    csr = dfq.Csr.rmat(15, 16, seed=9)
    srcs = csr.pick_sources(16, seed=4)
    with dfq.DeviceGraph(csr) as g, dfq.DeviceBuffer(4 * csr.n * 16) as buf:
        g.set_config(direction_opt=1)
        for rep in range(2):
            g.bfs_device_multi(srcs, buf.ptr)
            for i, s in enumerate(srcs):
                np.testing.assert_array_equal(buf.to_numpy(count=csr.n, offset=i * csr.n),
                                              oracle.bfs(csr.row_offsets, csr.col_idx, int(s)))
            r = g.bfs(int(srcs[rep]))
            np.testing.assert_array_equal(r.level, oracle.bfs(csr.row_offsets, csr.col_idx, int(srcs[rep])))
