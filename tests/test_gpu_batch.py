"""Batched searches (include/dfq_batch.h): every query's levels equal the
oracle's and the plain engine's, whatever tier and plan runs it."""
import numpy as np
import pytest

from conftest import UNREACHED

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make_graphs(dfq):
    gs = [dfq.Csr.rmat(s, 8, seed=100 + s) for s in (6, 8, 10, 11, 12, 14, 16)]
    gs.append(dfq.Csr.amazon_like(5000, seed=9))           # directed
    gs.append(dfq.Csr.grid2d(120, 90, 0.05, seed=3))        # long diameter
    gs.append(dfq.Csr.rmat(17, 8, seed=117))                # 131072 vertices: largest wg-tier graph
    gs.append(dfq.Csr.rmat(18, 8, seed=118))                # grid tier only
    return gs


def queries(gs, per=3):
    gi, so = [], []
    for i, g in enumerate(gs):
        for s in g.pick_sources(per, seed=i):
            gi.append(i)
            so.append(int(s))
    return np.array(gi), np.array(so)


@pytest.fixture(scope="module")
def batch_setup(dfq, oracle):
    gs = make_graphs(dfq)
    gi, so = queries(gs)
    expect = [oracle.bfs(gs[g].row_offsets, gs[g].col_idx, int(s)) for g, s in zip(gi, so)]
    return gs, gi, so, expect


def check_all(b, gi, so, expect):
    for q in range(len(gi)):
        lv = b.levels(q)
        np.testing.assert_array_equal(lv, expect[q], err_msg=f"query {q} graph {gi[q]} source {so[q]}")
        depth, reached = b.result(q)
        e = expect[q]
        assert reached == int((e != UNREACHED).sum())
        assert depth == int(e[e != UNREACHED].max()) + 1


@pytest.mark.parametrize("mode", ["default", "all_wg", "all_grid", "no_direction", "block512", "teams1",
                                  "all_grid_teams2"])
def test_batch_levels_match_oracle(dfq, batch_setup, mode):
    gs, gi, so, expect = batch_setup
    with dfq.Batch(gs) as b:
        b.set_queries(gi, so)
        for k in range(32):
            if mode == "all_wg" and (1 << k) <= dfq.WG_MAX_VERTICES:
                b.set_plan(k, tier=dfq.TIER_WORKGROUP)
            if mode in ("all_grid", "all_grid_teams2"):
                b.set_plan(k, tier=dfq.TIER_GRID)
            if mode == "no_direction":
                b.set_plan(k, direction_opt=0)
        if mode == "block512":
            b.set_block_size(512)
        if mode == "teams1":
            b.set_team_xcds(1)  # grid-tier queries on one-XCD teams, several at once
        if mode == "all_grid_teams2":
            b.set_team_xcds(2)
        for rep in range(3):  # the first run captures the graph, the others replay it
            say("batch", mode, "run", rep)
            st = b.run()
            assert st.queries == len(gi)
            check_all(b, gi, so, expect)
        if mode in ("all_grid", "all_grid_teams2"):
            assert st.wg_queries == 0
        assert st.team_xcds == {"teams1": 1, "all_grid_teams2": 2}.get(mode, 0)
        if mode == "all_wg":
            assert st.grid_queries == 3  # only the scale-18 graph


def test_batch_matches_plain_engine(dfq, batch_setup):
    gs, gi, so, expect = batch_setup
    with dfq.Batch(gs) as b:
        b.set_queries(gi, so)
        b.run()
        for q in range(0, len(gi), 4):
            with dfq.DeviceGraph(gs[gi[q]]) as g:
                np.testing.assert_array_equal(g.bfs(int(so[q])).level, b.levels(q))


def test_batch_autotune_keeps_results(dfq, batch_setup):
    gs, gi, so, expect = batch_setup
    with dfq.Batch(gs) as b:
        b.set_queries(gi, so)
        say("autotune")
        rep = b.autotune(repeats=1)
        assert rep["candidates"] > 0 and rep["block_size"] in (512, 1024) and rep["team_xcds"] in (0, 1, 2, 4)
        buckets = {e["bucket"] for e in rep["buckets"]}
        assert 18 in buckets and rep["buckets"][-1]["plan"]["tier"] == dfq.TIER_GRID
        b.run()
        check_all(b, gi, so, expect)


def test_batch_requery_and_errors(dfq, batch_setup):
    gs, gi, so, expect = batch_setup
    with dfq.Batch(gs[:3]) as b:
        with pytest.raises(dfq.DfqError):
            b.run()  # no queries yet
        with pytest.raises(dfq.DfqError) as e:
            b.set_queries([0], [gs[0].n])
        assert e.value.code == dfq.ERR_RANGE
        b.set_queries([2, 0], [5, 1])
        b.run()
        for q, (g, s) in enumerate(((2, 5), (0, 1))):
            with dfq.DeviceGraph(gs[g]) as dg:
                np.testing.assert_array_equal(b.levels(q), dg.bfs(s).level)


def test_batch_config4_full_workload(dfq, oracle):
    """Config 4 at its stated size, built exactly as bench.py builds it: 512
    undirected RMAT graphs (edgefactor 8, scales 10..20, p(n) ~ n^-1.5), 4
    sources each, 2,048 queries in one launch.  After autotuning, the first
    run captures the hipGraph and every later run replays it; every query's
    levels equal the plain engine's bit for bit, and a sample equals the
    oracle's."""
    import bench
    scales = bench.batch_scales(512, seed=4)
    gs = [dfq.Csr.rmat(int(s), 8, 0.57, 0.19, 0.19, seed=1000 + i) for i, s in enumerate(scales)]
    gi, so = [], []
    for i, g in enumerate(gs):
        for s in g.pick_sources(4, seed=i):
            gi.append(i)
            so.append(int(s))
    gi, so = np.array(gi), np.array(so)
    assert len(gi) == 2048 and min(scales) >= 10 and max(scales) <= 20
    say("config 4:", len(gs), "graphs,", sum(g.n for g in gs), "vertices,", len(gi), "queries")
    with dfq.Batch(gs) as b:
        b.set_queries(gi, so)
        tune = b.autotune(repeats=1)
        say("autotune: block", tune["block_size"], "team xcds", tune["team_xcds"], "buckets", len(tune["buckets"]))
        runs = [b.run() for _ in range(3)]
        assert all(st.queries == 2048 for st in runs)
        assert [bool(st.graph_replay) for st in runs[1:]] == [True, True], \
            [st.graph_replay for st in runs]
        by_graph = {}
        for q in range(len(gi)):
            by_graph.setdefault(int(gi[q]), []).append(q)
        checked = 0
        for g, qs in by_graph.items():
            with dfq.DeviceGraph(gs[g]) as dg:
                dg.set_config(direction_opt=1)
                for q in qs:
                    ref = dg.bfs(int(so[q]))
                    lv = b.levels(q)
                    np.testing.assert_array_equal(lv, ref.level, err_msg=f"query {q} graph {g}")
                    depth, reached = b.result(q)
                    assert (depth, reached) == (ref.depth, ref.reached)
                    if q % 32 == 0:
                        np.testing.assert_array_equal(
                            lv, oracle.bfs(gs[g].row_offsets, gs[g].col_idx, int(so[q])),
                            err_msg=f"oracle: query {q} graph {g}")
                        checked += 1
        say("all 2048 queries equal the plain engine;", checked, "equal the oracle")
