"""Served searches (include/dfq_serve.h): one persistent launch, continuous
batching.  Every query's levels equal the oracle's, and a query gives the
same levels alone as in the mix of a trace where many run at once."""
import numpy as np
import pytest

from conftest import UNREACHED

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make_graphs(dfq):
    return [
        dfq.Csr.rmat(12, 16, seed=212),           # symmetric, bottom-up levels
        dfq.Csr.rmat(15, 16, seed=215),
        dfq.Csr.rmat(17, 8, seed=317),
        dfq.Csr.amazon_like(20000, seed=7),       # directed: top-down only
        dfq.Csr.grid2d(200, 150, 0.03, seed=5),   # several hundred levels
    ]


@pytest.fixture(scope="module")
def serve_setup(dfq, oracle):
    gs = make_graphs(dfq)
    rng = np.random.default_rng(55)
    nq = 160
    gi = rng.integers(0, len(gs), nq).astype(np.uint32)
    so = np.array([int(gs[g].pick_sources(1, seed=1000 + q)[0]) for q, g in enumerate(gi)], np.uint32)
    # Poisson arrivals, 20 us apart on average: far more than one query in flight
    arrival = np.cumsum(rng.exponential(20.0, nq))
    expect = {}
    for g, s in zip(gi, so):
        if (g, s) not in expect:
            expect[(g, s)] = oracle.bfs(gs[g].row_offsets, gs[g].col_idx, int(s))
    return gs, gi, so, arrival, expect


def check_record(rec, e, g):
    assert rec["reached"] == int((e != UNREACHED).sum())
    assert rec["depth"] == int(e[e != UNREACHED].max()) + 1
    # traversed edges: out-degree sum of the reached vertices
    assert rec["edges"] == int(g.degrees()[e != UNREACHED].sum())


def max_overlap(recs):
    ev = [(r["admit_us"], 1) for r in recs] + [(r["done_us"], -1) for r in recs]
    cur = best = 0
    for _, d in sorted(ev, key=lambda x: (x[0], x[1])):
        cur += d
        best = max(best, cur)
    return best


def test_serve_mix_equals_alone_and_oracle(dfq, serve_setup):
    gs, gi, so, arrival, expect = serve_setup
    with dfq.Server(gs, xcds_per_team=1) as srv:
        srv.start()
        info = srv.info()
        say("server", info)
        assert info["running"] == 1 and info["teams"] >= 1
        assert sum(info["team_blocks"]) == info["grid_blocks"]
        say("replay", len(gi), "queries")
        recs, outs = srv.replay(gi, so, arrival, keep_levels=True)
        for q in range(len(gi)):
            e = expect[(gi[q], so[q])]
            np.testing.assert_array_equal(outs[q], e, err_msg=f"mix: query {q} graph {gi[q]} source {so[q]}")
            check_record(recs[q], e, gs[gi[q]])
        assert (recs["done_us"] >= recs["admit_us"]).all() and (recs["admit_us"] >= recs["arrival_us"]).all()
        say("max in flight", max_overlap(recs), "teams used", sorted(set(recs["team"].tolist())))
        if info["teams"] > 1:
            assert max_overlap(recs) > 1
            assert len(set(recs["team"].tolist())) > 1
        # the same queries one at a time
        for q in range(0, len(gi), 4):
            t = srv.submit(int(gi[q]), int(so[q]))
            r = srv.wait(t)
            lv = srv.levels(t, int(gi[q]))
            srv.release(t)
            np.testing.assert_array_equal(lv, outs[q], err_msg=f"alone: query {q}")
            assert r.reached == recs[q]["reached"] and r.depth == recs[q]["depth"]
        assert srv.info()["in_flight"] == 0
        srv.stop()
        assert srv.info()["running"] == 0


@pytest.mark.parametrize("xpt,block", [(2, 1024), (4, 1024), (0, 1024), (1, 512)])
def test_serve_team_shapes(dfq, serve_setup, xpt, block):
    gs, gi, so, arrival, expect = serve_setup
    sel = slice(0, 60)
    with dfq.Server(gs, xcds_per_team=xpt, block_size=block) as srv:
        srv.start()
        say("shape", xpt, block, srv.info())
        recs, outs = srv.replay(gi[sel], so[sel], arrival[sel], keep_levels=True)
        for q in range(len(recs)):
            e = expect[(gi[q], so[q])]
            np.testing.assert_array_equal(outs[q], e, err_msg=f"xpt {xpt}: query {q} graph {gi[q]}")
            check_record(recs[q], e, gs[gi[q]])
        srv.stop()


def test_serve_parents_and_direction_plans(dfq, serve_setup):
    gs, gi, so, arrival, expect = serve_setup
    with dfq.Server(gs) as srv:
        srv.set_plan(1, direction_opt=0, small_max=16, medium_max=64, large_max=1024)
        srv.start()
        tickets = []
        for q in range(0, 40, 2):
            tickets.append((q, srv.submit(int(gi[q]), int(so[q]), dfq.SERVE_PARENTS)))
        for q, t in tickets:
            srv.wait(t)
            lv, pa = srv.levels(t, int(gi[q]), parents=True)
            srv.release(t)
            np.testing.assert_array_equal(lv, expect[(gi[q], so[q])], err_msg=f"query {q}")
            rep = gs[gi[q]].validate(int(so[q]), lv, pa)
            assert rep["errors"] == 0, (q, rep)
        srv.stop()


def test_serve_admission_by_footprint(dfq, oracle, serve_setup):
    gs, gi, so, arrival, expect = serve_setup
    fp = [dfq.serve_footprint(g.n, g.m, 256, 1, 0) for g in gs]
    big = int(np.argmax(fp))
    small = int(np.argmin(fp))
    # room for one working set of the largest graph only
    arena = fp[big] + fp[small] // 2
    with dfq.Server(gs, arena_bytes=arena) as srv:
        srv.start()
        info = srv.info()
        assert info["arena_bytes"] >= arena
        s_big = int(gs[big].pick_sources(1, seed=1)[0])
        t1 = srv.submit(big, s_big)
        with pytest.raises(dfq.DfqError) as ei:
            srv.submit(big, s_big)        # held range leaves no room until released
        assert ei.value.code == dfq.ERR_BUSY
        srv.wait(t1)
        with pytest.raises(dfq.DfqError) as ei:
            srv.submit(big, s_big)        # finished but not released: still held
        assert ei.value.code == dfq.ERR_BUSY
        np.testing.assert_array_equal(srv.levels(t1, big),
                                      oracle.bfs(gs[big].row_offsets, gs[big].col_idx, s_big))
        srv.release(t1)
        t2 = srv.submit(big, s_big)
        srv.wait(t2)
        srv.release(t2)
        # the trace still completes, one large working set at a time
        recs, outs = srv.replay(gi[:50], so[:50], np.zeros(50), keep_levels=True)
        for q in range(50):
            np.testing.assert_array_equal(outs[q], expect[(gi[q], so[q])], err_msg=f"query {q}")
        srv.stop()
    # a working set larger than the whole arena is refused outright
    with dfq.Server(gs, arena_bytes=fp[small] * 2) as srv:
        srv.start()
        with pytest.raises(dfq.DfqError) as ei:
            srv.submit(big, 0)
        assert ei.value.code == dfq.ERR_CAPACITY
        srv.stop()


def test_serve_restart_and_results_after_stop(dfq, serve_setup):
    gs, gi, so, arrival, expect = serve_setup
    with dfq.Server(gs, ring_slots=8) as srv:
        for rnd in range(3):
            srv.start()
            say("restart round", rnd, srv.info()["teams"])
            # more queries than the ring holds: slots are reused
            recs, outs = srv.replay(gi[:20], so[:20], np.zeros(20), keep_levels=True)
            for q in range(20):
                np.testing.assert_array_equal(outs[q], expect[(gi[q], so[q])], err_msg=f"round {rnd} q {q}")
            held = [(q, srv.submit(int(gi[q]), int(so[q]))) for q in range(rnd * 5, rnd * 5 + 5)]
            for q, t in held:
                srv.wait(t)
            srv.stop()
            # finished queries stay readable after the launch ended
            for q, t in held:
                assert srv.poll(t) is not None
                np.testing.assert_array_equal(srv.levels(t, int(gi[q])), expect[(gi[q], so[q])])
                srv.release(t)
            assert srv.info()["in_flight"] == 0


@pytest.fixture(scope="module")
def config5_setup(dfq):
    """Config 5 graphs at their stated sizes, built exactly as bench.py
    builds them: 8 resident undirected RMAT graphs of scales 18, 19, 20, 21,
    22, 22, 23, 24 (edgefactor 16), queries picking graph i with p ~ 1/scale_i
    and a uniform degree>0 source.  A trace of 3,000 Poisson arrivals at
    about half the served rate (many queries in flight) and a dense sample of
    48 queries, with the plain engine's results: reached count and depth of
    every query of the trace, full level arrays of the sample."""
    import bench
    scales = bench.SERVE_SCALES
    gs = [dfq.Csr.rmat(s, 16, 0.57, 0.19, 0.19, seed=5000 + i) for i, s in enumerate(scales)]
    say("config 5:", sum(g.n for g in gs), "vertices,", sum(g.m for g in gs), "entries")
    gi, so, arr = bench.serve_trace(gs, scales, 3000, 4000.0, seed=13)
    g2, s2, _ = bench.serve_trace(gs, scales, 48, None, seed=17)
    arr2 = np.arange(48) * 5.0  # 5 us apart: the whole sample in flight together
    # the plain engine, one launch per query
    plain = np.zeros((len(gi), 2), np.uint64)
    plain_lv = [None] * len(g2)
    with dfq.DeviceBuffer(4 * max(g.n for g in gs)) as buf:
        for g in range(len(gs)):
            with dfq.DeviceGraph(gs[g]) as dg:
                dg.set_config(direction_opt=1)
                for q in np.flatnonzero(gi == g):
                    st = dg.bfs_device(int(so[q]), buf.ptr)
                    plain[q] = (st.reached, st.depth)
                for q in np.flatnonzero(g2 == g):
                    plain_lv[q] = dg.bfs(int(s2[q])).level
    assert all(lv is not None for lv in plain_lv)
    return gs, (gi, so, arr), (g2, s2, arr2), plain, plain_lv


def test_serve_config5_resident_graphs(dfq, oracle, config5_setup):
    """Config 5, one search per task: every query of the Poisson trace has
    the plain engine's reached count and depth, the dense sample gives the
    plain engine's full level arrays, and one query per graph the oracle's."""
    gs, (gi, so, arr), (g2, s2, arr2), plain, plain_lv = config5_setup
    with dfq.Server(gs, xcds_per_team=1) as srv:
        srv.start()
        say("server", srv.info())
        recs, _ = srv.replay(gi, so, arr)
        say("3000 Poisson queries: max in flight", max_overlap(recs))
        assert max_overlap(recs) > 1
        np.testing.assert_array_equal(recs["reached"], plain[:, 0])
        np.testing.assert_array_equal(recs["depth"].astype(np.uint64), plain[:, 1])
        recs2, outs = srv.replay(g2, s2, arr2, keep_levels=True)
        say("48-query mix: max in flight", max_overlap(recs2))
        srv.stop()
    seen = set()
    for q in range(len(g2)):
        np.testing.assert_array_equal(outs[q], plain_lv[q], err_msg=f"query {q} graph {g2[q]}")
        if int(g2[q]) not in seen:
            seen.add(int(g2[q]))
            g = gs[int(g2[q])]
            say("oracle, graph", int(g2[q]), "source", int(s2[q]))
            np.testing.assert_array_equal(outs[q], oracle.bfs(g.row_offsets, g.col_idx, int(s2[q])),
                                          err_msg=f"oracle: query {q} graph {g2[q]}")
    say("graphs checked against the oracle:", sorted(seen))


def test_serve_config5_group_tasks(dfq, config5_setup):
    """Config 5 through group tasks on one-XCD teams (the served form
    bench.py measures): the same trace and sample give the plain engine's
    results, the sample's full level arrays included."""
    gs, (gi, so, arr), (g2, s2, arr2), plain, plain_lv = config5_setup
    with dfq.Server(gs, xcds_per_team=1, groups=True) as srv:
        srv.start()
        say("group server", srv.info())
        recs, _ = srv.replay(gi, so, arr)
        np.testing.assert_array_equal(recs["reached"], plain[:, 0])
        np.testing.assert_array_equal(recs["depth"].astype(np.uint64), plain[:, 1])
        _, outs = srv.replay(g2, s2, arr2, keep_levels=True)
        srv.stop()
    for q in range(len(g2)):
        np.testing.assert_array_equal(outs[q], plain_lv[q], err_msg=f"group tasks: query {q} graph {g2[q]}")


# ---- group mode: queries of one graph share traversals --------------------

@pytest.mark.parametrize("xpt,block", [(1, 1024), (1, 512), (2, 1024), (0, 1024)])
def test_serve_groups_mix_equals_oracle(dfq, serve_setup, xpt, block):
    """Group-mode server (bit-parallel multi-source group tasks): the
    Poisson mix of the four graphs whose depth fits the group planes; every
    query's levels, reach, depth and traversed entries equal the oracle's,
    and queries did share traversals (fewer tasks than queries)."""
    gs, gi, so, arrival, expect = serve_setup
    keep = gi < 4  # the 200 x 150 grid is deeper than 255 levels
    gi2, so2, ar2 = gi[keep], so[keep], arrival[keep]
    with dfq.Server(gs, xcds_per_team=xpt, block_size=block, groups=True) as srv:
        srv.start()
        say("group server", srv.info())
        recs, outs = srv.replay(gi2, so2, ar2, keep_levels=True)
        for q in range(len(gi2)):
            e = expect[(gi2[q], so2[q])]
            np.testing.assert_array_equal(outs[q], e, err_msg=f"groups xpt {xpt}: query {q} graph {gi2[q]}")
            check_record(recs[q], e, gs[gi2[q]])
        tasks = len(set(zip(recs["admit_us"].tolist(), recs["done_us"].tolist())))
        say("queries", len(gi2), "group tasks", tasks, "max in flight", max_overlap(recs))
        assert tasks < len(gi2)
        srv.stop()


def test_serve_groups_direct_api(dfq, oracle, serve_setup):
    """submit_group: per-member tickets, results and levels; repeated
    sources; a group of one through submit; release frees the range with
    the last member; the deep grid is refused with ERR_CAPACITY."""
    gs, gi, so, arrival, expect = serve_setup
    g = 1
    srcs = list(gs[g].pick_sources(40, seed=8)) + [int(gs[g].pick_sources(1, seed=8)[0])] * 3
    ref = {int(s): oracle.bfs(gs[g].row_offsets, gs[g].col_idx, int(s)) for s in srcs}
    with dfq.Server(gs, groups=True) as srv:
        srv.start()
        tks = srv.submit_group(g, srcs)
        assert len(tks) == len(srcs) and len(set(tks)) == len(tks)
        t1 = srv.submit(0, int(gs[0].pick_sources(1, seed=2)[0]))
        for s, t in zip(srcs, tks):
            r = srv.wait(t)
            e = ref[int(s)]
            assert r.reached == int((e != UNREACHED).sum()) and r.depth == int(e[e != UNREACHED].max()) + 1
            np.testing.assert_array_equal(srv.levels(t, g), e)
        used = srv.info()["arena_used"]
        for t in tks[:-1]:
            srv.release(t)
        assert srv.info()["arena_used"] == used  # the last member still holds the range
        srv.release(tks[-1])
        assert srv.info()["arena_used"] < used
        srv.wait(t1)
        s0 = int(gs[0].pick_sources(1, seed=2)[0])
        np.testing.assert_array_equal(srv.levels(t1, 0), oracle.bfs(gs[0].row_offsets, gs[0].col_idx, s0))
        srv.release(t1)
        # 4: the 200 x 150 grid from a corner: past the 255 levels of the planes
        t2 = srv.submit(4, 0)
        with pytest.raises(dfq.DfqError) as e:
            srv.wait(t2)
        assert e.value.code == dfq.ERR_CAPACITY
        srv.release(t2)
        with pytest.raises(dfq.DfqError):
            srv.submit(0, 0, dfq.SERVE_PARENTS)  # no parents in group mode
        assert srv.info()["in_flight"] == 0
        srv.stop()
