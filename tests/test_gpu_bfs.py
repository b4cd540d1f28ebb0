"""Parity of the HIP engine (through the C ABI) with the oracle and the
golden vectors: levels bit-exact for every vertex, parent tree valid."""
import numpy as np
import pytest

from conftest import UNREACHED, csr_from_edges_np, golden_names, load_golden

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def check_result(dfq, csr, source, res, expect):
    np.testing.assert_array_equal(res.level, expect, err_msg=f"source {source}")
    if res.parent is not None:
        rep = csr.validate(source, res.level, res.parent)
        assert rep["errors"] == 0, rep
    reached = int((expect != UNREACHED).sum())
    assert res.reached == reached
    assert res.depth == int(expect[expect != UNREACHED].max()) + 1
    # E_r, the numerator of TEPS: CSR entries of the reached vertices
    assert res.edges == int(np.diff(csr.row_offsets)[expect != UNREACHED].sum())


def test_device_info(dfq):
    info = dfq.device_info()
    say("device", info)
    assert info["cu_count"] > 0 and info["cooperative"] == 1
    if info["arch"].startswith("gfx950"):
        assert info["cu_count"] == 256 and info["xcc_count"] == 8


@pytest.mark.parametrize("name", golden_names())
def test_golden(dfq, oracle, name):
    gd = load_golden(name)
    csr = dfq.Csr.from_arrays(gd["row_offsets"], gd["col_idx"])
    with dfq.DeviceGraph(csr) as g:
        for s, expect in zip(gd["sources"], gd["levels"]):
            say("golden", name, "source", int(s))
            res = g.bfs(int(s), parents=True)
            check_result(dfq, csr, int(s), res, expect)
            np.testing.assert_array_equal(expect, oracle.bfs(gd["row_offsets"], gd["col_idx"], int(s)))


CONFIGS = [
    dict(block_size=256, small_max=32, medium_max=2048, large_max=32768),
    dict(block_size=512, small_max=32, medium_max=2048, large_max=32768),
    dict(block_size=256, small_max=1, medium_max=1, large_max=1),      # everything huge
    dict(block_size=256, small_max=64, medium_max=64, large_max=64),   # no wave/WG bins
    dict(block_size=256, small_max=0, medium_max=4096, large_max=4096),
    dict(block_size=512, small_max=8, medium_max=64, large_max=1024),
    dict(block_size=256, small_max=16, medium_max=256, large_max=16384, blocks_per_cu=1),
    dict(block_size=1024, small_max=32, medium_max=2048, large_max=32768),
    dict(block_size=1024, small_max=4, medium_max=128, large_max=512),
]


@pytest.mark.parametrize("ci", range(len(CONFIGS)))
def test_bin_configs_on_hubs(dfq, oracle, ci):
    gd = load_golden("bin_hubs_50000")
    csr = dfq.Csr.from_arrays(gd["row_offsets"], gd["col_idx"])
    with dfq.DeviceGraph(csr) as g:
        g.set_config(**CONFIGS[ci])
        for s, expect in zip(gd["sources"], gd["levels"]):
            say("config", CONFIGS[ci], "source", int(s))
            check_result(dfq, csr, int(s), g.bfs(int(s), parents=True), expect)


def test_closed_forms(dfq):
    # long directed path: one vertex per level (2000 levels, 2000 barriers)
    n = 2000
    ro, ci = csr_from_edges_np(n, np.arange(n - 1), np.arange(1, n))
    csr = dfq.Csr.from_arrays(ro, ci)
    with dfq.DeviceGraph(csr) as g:
        say("path 2000")
        r = g.bfs(0, parents=True)
        np.testing.assert_array_equal(r.level, np.arange(n, dtype=np.uint32))
        np.testing.assert_array_equal(r.parent[1:], np.arange(n - 1, dtype=np.uint32))
        r = g.bfs(n - 1)
        assert r.level[n - 1] == 0 and (r.level[:-1] == UNREACHED).all()
    # star with a 300k-degree hub: whole-grid expansion of a single vertex
    n = 300001
    ro, ci = csr_from_edges_np(n, np.zeros(n - 1, int), np.arange(1, n))
    csr = dfq.Csr.from_arrays(ro, ci)
    with dfq.DeviceGraph(csr) as g:
        say("star 300k")
        r = g.bfs(0, parents=True)
        assert r.level[0] == 0 and (r.level[1:] == 1).all() and (r.parent[1:] == 0).all()
    # complete binary tree
    n = (1 << 16) - 1
    child = np.arange(1, n)
    ro, ci = csr_from_edges_np(n, (child - 1) // 2, child)
    csr = dfq.Csr.from_arrays(ro, ci)
    with dfq.DeviceGraph(csr) as g:
        say("binary tree")
        r = g.bfs(0, parents=True)
        np.testing.assert_array_equal(r.level, np.floor(np.log2(np.arange(n) + 1)).astype(np.uint32))
        assert csr.validate(0, r.level, r.parent)["errors"] == 0


@pytest.mark.parametrize("direction_opt", [0, 1])
def test_huge_hubs_closed_form(dfq, direction_opt):
    """Undirected hubs at the reference's largest bin edges and beyond:
    degree 2^20 + 1 (vertex 0, joined to 1..2^20+1), 167,937 (vertex 1),
    167,936 (vertex 2), 16,385 (vertex 3) and 16,384 (vertex 4), each of
    the last four with leaves of its own.  Closed form from vertex 0: hub
    0, its neighbours 1, the other hubs' own leaves 2; from vertex 1: 1 is
    0, vertex 0 and 1's leaves 1, the rest 2, the other hubs' leaves 3."""
    big = (1 << 20) + 1
    degs = [167937, 167936, 16385, 16384]
    src, dst, lvl0 = [], [], []
    src.append(np.zeros(big, np.int64))
    dst.append(np.arange(1, big + 1, dtype=np.int64))
    nxt = big + 1
    own = []
    for h, d in zip([1, 2, 3, 4], degs):
        leaves = np.arange(nxt, nxt + d, dtype=np.int64)
        own.append(leaves)
        src.append(np.full(d, h, np.int64))
        dst.append(leaves)
        nxt += d
    n = nxt
    s = np.concatenate(src)
    t = np.concatenate(dst)
    ro, ci = csr_from_edges_np(n, np.concatenate([s, t]), np.concatenate([t, s]))
    csr = dfq.Csr.from_arrays(ro, ci)
    e0 = np.full(n, 1, np.uint32)
    e0[0] = 0
    for leaves in own:
        e0[leaves] = 2
    e1 = np.full(n, 2, np.uint32)
    e1[1] = 0
    e1[0] = 1
    e1[own[0]] = 1
    for leaves in own[1:]:
        e1[leaves] = 3
    with dfq.DeviceGraph(csr) as g:
        g.set_config(direction_opt=direction_opt)
        say("huge hubs", direction_opt)
        r = g.bfs(0, parents=True)
        np.testing.assert_array_equal(r.level, e0)
        assert csr.validate(0, r.level, r.parent)["errors"] == 0
        r = g.bfs(1, parents=True)
        np.testing.assert_array_equal(r.level, e1)
        assert csr.validate(1, r.level, r.parent)["errors"] == 0
        lv = dfq.DeviceBuffer(4 * n * 2)
        st, _ = g.bfs_device_seq([0, 1], lv.ptr, n)
        both = lv.to_numpy().reshape(2, n)
        np.testing.assert_array_equal(both[0], e0)
        np.testing.assert_array_equal(both[1], e1)


def test_single_vertex_and_isolated_source(dfq):
    csr = dfq.Csr.from_arrays([0, 0], np.zeros(0, np.uint32))
    with dfq.DeviceGraph(csr) as g:
        r = g.bfs(0, parents=True)
        assert r.level.tolist() == [0] and r.parent.tolist() == [0] and r.depth == 1
    ro, ci = csr_from_edges_np(5, [1, 2], [2, 3])
    csr = dfq.Csr.from_arrays(ro, ci)
    with dfq.DeviceGraph(csr) as g:
        r = g.bfs(4)
        assert r.level.tolist() == [UNREACHED] * 4 + [0]
        with pytest.raises(dfq.DfqError) as e:
            g.bfs(5)
        assert e.value.code == dfq.ERR_RANGE


@pytest.mark.parametrize("scale", [10, 14, 17])
def test_rmat_many_sources(dfq, oracle, scale):
    csr = dfq.Csr.rmat(scale, 16, seed=scale)
    srcs = csr.pick_sources(8, seed=7)
    with dfq.DeviceGraph(csr) as g:
        for s in srcs:
            say("rmat", scale, "source", int(s))
            check_result(dfq, csr, int(s), g.bfs(int(s), parents=True),
                         oracle.bfs(csr.row_offsets, csr.col_idx, int(s)))


DO_CONFIGS = [
    dict(block_size=256, direction_opt=1, bu_enter_div=64, bu_exit_div=512),
    dict(block_size=1024, direction_opt=1, bu_enter_div=256, bu_exit_div=1024),
    dict(block_size=512, direction_opt=1, bu_enter_div=1000000, bu_exit_div=1),  # bottom-up from level 1
]


@pytest.mark.parametrize("ci", range(len(DO_CONFIGS)))
def test_direction_optimizing_levels_match(dfq, oracle, ci):
    """Bottom-up levels on symmetric graphs give the same levels."""
    for scale, ef in ((12, 8), (16, 16)):
        csr = dfq.Csr.rmat(scale, ef, seed=100 + scale)
        with dfq.DeviceGraph(csr) as g:
            assert g.symmetric
            g.set_config(**DO_CONFIGS[ci])
            for s in csr.pick_sources(4, seed=ci):
                say("DO", DO_CONFIGS[ci], "rmat", scale, "source", int(s))
                r = g.bfs(int(s), parents=True)
                check_result(dfq, csr, int(s), r, oracle.bfs(csr.row_offsets, csr.col_idx, int(s)))
                assert r.bottom_up_levels > 0
    grid = dfq.Csr.grid2d(300, 200, 0.05, seed=5)
    with dfq.DeviceGraph(grid) as g:
        g.set_config(**DO_CONFIGS[ci])
        say("DO grid")
        check_result(dfq, grid, 0, g.bfs(0, parents=True), oracle.bfs(grid.row_offsets, grid.col_idx, 0))


def test_directed_graph_never_runs_bottom_up(dfq, oracle):
    csr = dfq.Csr.amazon_like(50000, seed=3)
    with dfq.DeviceGraph(csr) as g:
        assert not g.symmetric
        g.set_config(direction_opt=1, bu_enter_div=1000000, bu_exit_div=1)
        r = g.bfs(0, parents=True)
        assert r.bottom_up_levels == 0
        check_result(dfq, csr, 0, r, oracle.bfs(csr.row_offsets, csr.col_idx, 0))


def test_amazon_like_config1(dfq, oracle):
    """Config 1 graph (262,144 vertices, directed), source 0."""
    csr = dfq.Csr.amazon_like(262144, seed=0xD0F1)
    with dfq.DeviceGraph(csr) as g:
        say("amazon-like source 0")
        check_result(dfq, csr, 0, g.bfs(0, parents=True), oracle.bfs(csr.row_offsets, csr.col_idx, 0))


def test_rmat22_full_size(dfq, oracle):
    """Config 2 graph at full size: RMAT scale 22, edgefactor 16."""
    csr = dfq.Csr.rmat(22, 16, seed=22)
    srcs = csr.pick_sources(64, seed=2022)
    with dfq.DeviceGraph(csr) as g:
        for i, s in enumerate(srcs[:4]):
            say("rmat22 source", int(s))
            check_result(dfq, csr, int(s), g.bfs(int(s), parents=(i % 2 == 0)),
                         oracle.bfs(csr.row_offsets, csr.col_idx, int(s)))
        # the rest: size-independent properties (certificate + reach count)
        e_ref = None
        for s in srcs[4:16]:
            say("rmat22 certificate source", int(s))
            r = g.bfs(int(s), parents=True)
            assert csr.validate(int(s), r.level, r.parent)["errors"] == 0
            e, v, _ = csr.reached_work(r.level)
            e_ref = e_ref or e
            assert e == e_ref  # every source lies in the one giant component


def test_rmat22_timed_kernel_bench_config(dfq, oracle):
    """Exactly what bench.py times: the sequence kernel (one launch, 64
    searches, search i writing at level + i * n) in the hybrid config on the
    config-2 graph (RMAT-22, graph seed 22, sources seed 2022).  Eight
    searches bit-exact against the oracle, the other 56 certified; the
    parents of a second launch certified on all 64."""
    import bench
    csr = dfq.Csr.rmat(22, 16, seed=22)
    srcs = csr.pick_sources(64, seed=2022)
    n = csr.n
    with dfq.DeviceGraph(csr) as g, dfq.DeviceBuffer(4 * n * 64) as d_level, \
            dfq.DeviceBuffer(4 * n * 64) as d_parent:
        g.set_config(**bench.MODES["hybrid"])
        assert g.config().block_size == 1024 and g.config().direction_opt == 1
        say("rmat22 sequence launch, 64 sources, hybrid, stride n")
        stats, ms = g.bfs_device_seq(srcs, d_level.ptr, level_stride=n)
        assert ms > 0 and sum(st.bottom_up_levels for st in stats) > 0
        for i, s in enumerate(srcs):
            lv = d_level.to_numpy(count=n, offset=i * n)
            if i % 8 == 0:
                say("  oracle compare, search", i, "source", int(s))
                e = oracle.bfs(csr.row_offsets, csr.col_idx, int(s))
                np.testing.assert_array_equal(lv, e, err_msg=f"search {i} source {s}")
            else:
                assert csr.validate(int(s), lv)["errors"] == 0, (i, int(s))
            assert stats[i].reached == int((lv != UNREACHED).sum())
            assert stats[i].depth == int(lv[lv != UNREACHED].max()) + 1
        say("rmat22 sequence launch with parents")
        g.bfs_device_seq(srcs, d_level.ptr, level_stride=n, d_parent=d_parent.ptr, parent_stride=n)
        for i, s in enumerate(srcs):
            lv = d_level.to_numpy(count=n, offset=i * n)
            pa = d_parent.to_numpy(count=n, offset=i * n)
            assert csr.validate(int(s), lv, pa)["errors"] == 0, (i, int(s))


def test_grid_config3(dfq, oracle):
    """Config 3 graph: 4096 x 4096 grid, 2% edges dropped, source (0,0)."""
    csr = dfq.Csr.grid2d(4096, 4096, 0.02, seed=4096)
    with dfq.DeviceGraph(csr) as g:
        say("grid 4096")
        r = g.bfs(0, parents=True)
        check_result(dfq, csr, 0, r, oracle.bfs(csr.row_offsets, csr.col_idx, 0))
        assert r.depth > 8000


def test_repeat_runs_share_no_state(dfq, oracle):
    """Back-to-back searches on one resident graph: nothing left over."""
    csr = dfq.Csr.rmat(16, 8, seed=3)
    srcs = csr.pick_sources(6, seed=1)
    with dfq.DeviceGraph(csr) as g:
        expect = {int(s): oracle.bfs(csr.row_offsets, csr.col_idx, int(s)) for s in srcs}
        for rep in range(3):
            for s in srcs:
                r = g.bfs(int(s))
                np.testing.assert_array_equal(r.level, expect[int(s)], err_msg=f"rep {rep} src {s}")


def test_device_outputs(dfq, oracle):
    """dfq_bfs_device writes into caller-owned device buffers."""
    csr = dfq.Csr.rmat(15, 16, seed=9)
    s = int(csr.pick_sources(1, 3)[0])
    lv = dfq.DeviceBuffer(4 * csr.n)
    par = dfq.DeviceBuffer(4 * csr.n)
    with dfq.DeviceGraph(csr) as g:
        say("device outputs source", s)
        st = g.bfs_device(s, lv.ptr, par.ptr)
        assert st.kernel_ms > 0
        level, parent = lv.to_numpy(), par.to_numpy()
        np.testing.assert_array_equal(level, oracle.bfs(csr.row_offsets, csr.col_idx, s))
        assert csr.validate(s, level, parent)["errors"] == 0
        st = g.bfs_device(s, lv.ptr, None)  # parents optional
        np.testing.assert_array_equal(lv.to_numpy(), level)


@pytest.mark.parametrize("direction_opt", [0, 1])
def test_device_many_back_to_back(dfq, oracle, direction_opt):
    """dfq_bfs_device_many: per-search results equal single searches'; the
    output buffer holds the last source's levels."""
    csr = dfq.Csr.rmat(14, 16, seed=77)
    srcs = [int(s) for s in csr.pick_sources(12, 5)]
    lv = dfq.DeviceBuffer(4 * csr.n)
    with dfq.DeviceGraph(csr) as g:
        g.set_config(direction_opt=direction_opt)
        say("device many", direction_opt)
        stats = g.bfs_device_many(srcs, lv.ptr)
        assert len(stats) == len(srcs)
        for s, st in zip(srcs, stats):
            e = oracle.bfs(csr.row_offsets, csr.col_idx, s)
            assert st.kernel_ms > 0
            assert st.reached == int((e != UNREACHED).sum())
            assert st.depth == int(e[e != UNREACHED].max()) + 1
        np.testing.assert_array_equal(lv.to_numpy(), oracle.bfs(csr.row_offsets, csr.col_idx, srcs[-1]))
        with pytest.raises(dfq.DfqError):
            g.bfs_device_many([csr.n], lv.ptr)


SEQ_GRAPHS = [
    ("rmat14", lambda dfq: dfq.Csr.rmat(14, 16, seed=78)),
    ("grid256", lambda dfq: dfq.Csr.grid2d(256, 256, 0.05, seed=11)),   # long solo-level runs
    ("amazon", lambda dfq: dfq.Csr.amazon_like(65536, seed=5)),
]


@pytest.mark.parametrize("direction_opt", [0, 1])
@pytest.mark.parametrize("gname", [g[0] for g in SEQ_GRAPHS])
def test_device_seq_one_launch(dfq, oracle, gname, direction_opt):
    """dfq_bfs_device_seq: every search of one launch equals the oracle
    (levels kept apart with stride n, parents valid), stats match, and a
    stride of 0 leaves the last search's levels."""
    csr = dict(SEQ_GRAPHS)[gname](dfq)
    n = csr.n
    deg = np.diff(csr.row_offsets)
    srcs = [int(s) for s in csr.pick_sources(9, 17)]
    iso = np.flatnonzero(deg == 0)
    if iso.size:
        srcs.insert(3, int(iso[0]))         # isolated source mid-sequence
    srcs.append(srcs[0])                     # a repeated source
    cnt = len(srcs)
    lv = dfq.DeviceBuffer(4 * n * cnt)
    pa = dfq.DeviceBuffer(4 * n * cnt)
    with dfq.DeviceGraph(csr) as g:
        g.set_config(direction_opt=direction_opt)
        say("seq", gname, direction_opt, cnt)
        stats, launch_ms = g.bfs_device_seq(srcs, lv.ptr, n, pa.ptr, n)
        assert launch_ms > 0 and len(stats) == cnt
        levels = lv.to_numpy().reshape(cnt, n)
        parents = pa.to_numpy().reshape(cnt, n)
        for i, (s, st) in enumerate(zip(srcs, stats)):
            e = oracle.bfs(csr.row_offsets, csr.col_idx, s)
            np.testing.assert_array_equal(levels[i], e, err_msg=f"search {i} source {s}")
            rep = csr.validate(s, levels[i], parents[i])
            assert rep["errors"] == 0, (i, s, rep)
            assert st.reached == int((e != UNREACHED).sum())
            assert st.edges == int(deg[e != UNREACHED].sum()), (i, s)
            assert st.depth == int(e[e != UNREACHED].max()) + 1
            assert st.kernel_ms > 0
        assert sum(st.kernel_ms for st in stats) <= launch_ms * 1.05 + 0.05
        # stride 0: the searches overwrite one array; the last one stays
        one = dfq.DeviceBuffer(4 * n)
        g.bfs_device_seq(srcs[::-1], one.ptr)
        np.testing.assert_array_equal(one.to_numpy(), levels[0])
        # a plain launch afterwards shares no state with the sequence
        r = g.bfs(srcs[1])
        np.testing.assert_array_equal(r.level, levels[1])
        assert r.edges == stats[1].edges and r.reached == stats[1].reached


def test_device_seq_bad_input(dfq):
    csr = dfq.Csr.rmat(10, 8, seed=2)
    lv = dfq.DeviceBuffer(4 * csr.n * 2)
    with dfq.DeviceGraph(csr) as g:
        with pytest.raises(dfq.DfqError):
            g.bfs_device_seq([0, 1], lv.ptr, csr.n - 1)      # stride below n
        with pytest.raises(dfq.DfqError):
            g.bfs_device_seq([0, csr.n], lv.ptr, csr.n)      # source out of range
        st, ms = g.bfs_device_seq([], lv.ptr)
        assert st == [] and ms == 0.0


def test_profile_counters(dfq, oracle):
    """dfq_bfs_profile on a directed two-level graph whose second frontier
    (2,000 vertices, more than one per thread, so it goes through the bins)
    has out-degrees at every bin edge: per-bin counts, the closed-form idle
    lane slots of the wave, workgroup and work-item bins, entries, claims
    and frontiers; then level-by-level consistency on RMAT with bottom-up
    levels."""
    cfg = dict(block_size=1024, small_max=32, medium_max=256, large_max=4096, direction_opt=0)
    pattern = [1, 5, 32, 33, 64, 100, 256, 257, 1000, 1024, 4096, 4097, 10000]
    mids = 2000
    degs = np.array([pattern[i % len(pattern)] for i in range(mids)], np.int64)
    src = [np.zeros(mids, np.int64)]
    dst = [np.arange(1, mids + 1, dtype=np.int64)]
    nxt = mids + 1
    for i, d in enumerate(degs):
        src.append(np.full(d, i + 1, np.int64))
        dst.append(np.arange(nxt, nxt + d, dtype=np.int64))
        nxt += d
    n = nxt
    ro, ci = csr_from_edges_np(n, np.concatenate(src), np.concatenate(dst))
    csr = dfq.Csr.from_arrays(ro, ci)
    small = degs <= 32
    med = (degs > 32) & (degs <= 256)
    big = (degs > 256) & (degs <= 4096)
    huge = degs > 4096
    with dfq.DeviceGraph(csr) as g:
        g.set_config(**cfg)
        say("profile bins")
        rows, st = g.profile(0)
        assert st.depth == 3 and len(rows) == 3
        r0, r1 = rows[0], rows[1]
        assert r0["frontier"] == 1 and r0["claims"] == mids and r0["edges"] == mids
        assert r1["frontier"] == mids and r1["claims"] == int(degs.sum()) and r1["edges"] == int(degs.sum())
        assert rows[2]["frontier"] == int(degs.sum()) and rows[2]["claims"] == 0
        assert (r1["bin_small"], r1["bin_medium"], r1["bin_large"], r1["bin_huge"]) == \
            (int(small.sum()), int(med.sum()), int(big.sum()), int(huge.sum()))
        assert r1["idle_medium"] == int(((64 - degs[med] % 64) % 64).sum())
        assert r1["idle_large"] == int(((1024 - degs[big] % 1024) % 1024).sum())
        assert r1["idle_huge"] == int(((64 - degs[huge] % 64) % 64).sum())
        blocks = (mids + 63) // 64
        assert 0 <= r1["idle_small"] < 64 * blocks
        assert 0 < r1["edges_max_wg"] <= r1["edges"]
        assert r1["items"] == int(((degs[huge] + 255) // 256).sum())
        assert all(r["bottom_up"] == 0 and r["bu_candidates"] == 0 for r in rows)
        assert sum(r["claims"] for r in rows) == st.reached - 1

    csr = dfq.Csr.rmat(16, 16, seed=9)
    s = int(csr.pick_sources(1, 3)[0])
    e = oracle.bfs(csr.row_offsets, csr.col_idx, s)
    with dfq.DeviceGraph(csr) as g:
        g.set_config(direction_opt=1)
        say("profile rmat16")
        rows, st = g.profile(s)
        assert st.depth == len(rows) == int(e[e != UNREACHED].max()) + 1
        for L, r in enumerate(rows):
            assert r["frontier"] == int((e == L).sum())
            assert r["claims"] == int((e == L + 1).sum())
            if r["bottom_up"]:
                assert r["bu_candidates"] > 0
                assert r["bin_small"] + r["bin_medium"] + r["bin_large"] + r["bin_huge"] == 0
            else:
                assert r["bu_candidates"] == 0
        assert any(r["bottom_up"] for r in rows)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_launch_modes(dfq, oracle, mode):
    """Plain launches that assemble their grid (mode 0, the default),
    cooperative launches (1), and plain launches that give up at once and
    are run again cooperatively (2): the same levels and parents for single
    searches, back-to-back searches and the profiled search, and searches
    after a given-up launch are unaffected."""
    csr = dfq.Csr.rmat(14, 16, seed=31)
    srcs = [int(s) for s in csr.pick_sources(6, 7)]
    expect = {s: oracle.bfs(csr.row_offsets, csr.col_idx, s) for s in srcs}
    lv = dfq.DeviceBuffer(4 * csr.n)
    with dfq.DeviceGraph(csr) as g:
        g.set_config(direction_opt=1, launch_mode=mode)
        say("launch mode", mode)
        for s in srcs:
            r = g.bfs(s, parents=True)
            np.testing.assert_array_equal(r.level, expect[s])
            assert csr.validate(s, r.level, r.parent)["errors"] == 0
        stats = g.bfs_device_many(srcs, lv.ptr)
        for s, st in zip(srcs, stats):
            assert st.reached == int((expect[s] != UNREACHED).sum())
        np.testing.assert_array_equal(lv.to_numpy(), expect[srcs[-1]])
        rows, st = g.profile(srcs[0])
        assert sum(r["claims"] for r in rows) == st.reached - 1
        # alternate modes on the same graph: census words of every kind stay consistent
        g.set_config(launch_mode=(mode + 1) % 3)
        np.testing.assert_array_equal(g.bfs(srcs[1]).level, expect[srcs[1]])
        g.set_config(launch_mode=mode)
        np.testing.assert_array_equal(g.bfs(srcs[2]).level, expect[srcs[2]])
        st2, _ = g.bfs_device_seq(srcs, lv.ptr)
        np.testing.assert_array_equal(lv.to_numpy(), expect[srcs[-1]])


def test_launch_mode_rejected(dfq):
    csr = dfq.Csr.rmat(8, 4, seed=1)
    with dfq.DeviceGraph(csr) as g:
        with pytest.raises(dfq.DfqError):
            g.set_config(launch_mode=3)
