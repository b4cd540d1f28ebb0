#!/usr/bin/env python3
"""Benchmark of the MI355X DFQ-BFS engine on one GPU.

Workload (the headline configuration): undirected RMAT graph, scale 22
(4,194,304 vertices), edgefactor 16 (67,108,864 generated edges),
a/b/c/d = 0.57/0.19/0.19/0.05, self loops and duplicates removed,
symmetrised, vertex ids permuted, uint32 columns with int64 row offsets;
64 seeded sources drawn from vertices with degree > 0.

A step is one BFS from each of the 64 sources, all in ONE persistent
cooperative launch (dfq_bfs_device_seq: the grid runs the searches in turn,
outputs written to a device buffer that stays resident).  `value` is the
harmonic mean over the sources of the per-source GTEPS, GTEPS = (E_r / 2) /
t with E_r the directed CSR entries of the reached component (Graph500
convention for undirected graphs) and t the source's share of the launch:
its device-clock time inside the launch, scaled so that the 64 shares add
up to the hipEvent time of the whole launch (dispatch and kernel-start
census included), averaged over the timed steps.  The roofline entry
prices the algorithmic bytes of the step, the sum over the searches of
B_alg = 4 E_r + 16 V_r + 4 N (CSR entries, two int64 offsets per reached
vertex, the level array written in full), against 8 TB/s over the event
time of the launch.  "one_launch_per_search" repeats the measurement with
one launch per source (dfq_bfs_device_many: plain launches whose grid
assembles itself, each bracketed by its own hipEvents), the reference's
way of running a search.

cpu_baseline times the oracle (serial CPU BFS restating the reference's
host loop) on a bounded sample of the same sources and also checks that the
GPU levels equal the oracle's bit for bit on that sample.
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dfq-bfs_amd"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


MODES = {
    # every level top-down through the degree-binned frontier queues
    "topdown": dict(block_size=1024, small_max=32, medium_max=256, large_max=256,
                    direction_opt=0, bu_enter_div=40, bu_exit_div=96),
    # the same kernel with bottom-up levels for large frontiers
    "hybrid": dict(block_size=1024, small_max=32, medium_max=256, large_max=256,
                   direction_opt=1, bu_enter_div=40, bu_exit_div=96),
}


def run_mode(dfq, g, srcs, work, steps, warmup, d_level, launch="seq"):
    """Per-source mean ms over the timed steps, and wall ms per step.
    launch "seq": one launch per step, search i writing its levels at
    d_level + i * n (the output the bench checks afterwards), each source's
    device-clock time scaled so the step's shares add up to the launch's
    event time; "each": one launch per source, each bracketed by its own
    events."""
    def step():
        if launch == "seq":
            st, launch_ms = g.bfs_device_seq(srcs, d_level.ptr, level_stride=g.n)
            dev = np.array([x.kernel_ms for x in st])
            return dev * (launch_ms / dev.sum())
        return np.array([x.kernel_ms for x in g.bfs_device_many(srcs, d_level.ptr)])
    for _ in range(warmup):
        step()
    dfq.synchronize()
    times = np.zeros((steps, len(srcs)))
    t0 = time.perf_counter()
    for k in range(steps):
        times[k] = step()
    dfq.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    return dict(t=times.mean(axis=0), wall_ms_per_step=wall_ms)


def price(r, work, n):
    """GTEPS and roofline of per-source times r["t"] (ms) given each
    search's (E_r, V_r)."""
    t = r["t"]
    e_r = np.array([w[0] for w in work], dtype=np.float64)
    v_r = np.array([w[1] for w in work], dtype=np.float64)
    gteps = (e_r / 2) / (t * 1e-3) / 1e9
    hmean = len(gteps) / np.sum(1.0 / gteps)
    b_alg = 4 * e_r + 16 * v_r + 4 * n
    achieved = b_alg.sum() / (t.sum() * 1e-3) / 1e9
    return dict(gteps_hmean=float(hmean), gteps_median=float(np.median(gteps)),
                ms_per_bfs_mean=float(t.mean()), ms_per_bfs_median=float(np.median(t)),
                wall_ms_per_step=r["wall_ms_per_step"], achieved_gbs=float(achieved),
                b_alg_mean=float(b_alg.mean()), b_alg_step=float(b_alg.sum()),
                kernel_ms_total=float(t.sum()))


def host_cpu():
    """CPU model and logical core count of the machine running the bench."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(csr, srcs, levels_gpu, budget_s):
    """Serial oracle BFS on as many sources as fit in budget_s (at least 2)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import Oracle  # the checker; never the measured GPU path
    orc = Oracle()
    done, edges, secs, exact = 0, 0, 0.0, True
    for i, s in enumerate(srcs):
        t0 = time.perf_counter()
        lv = orc.bfs(csr.row_offsets, csr.col_idx, int(s))
        dt = time.perf_counter() - t0
        e, _, _ = csr.reached_work(lv)
        edges += e / 2
        secs += dt
        done += 1
        exact &= bool(np.array_equal(lv, levels_gpu[i]))
        if done >= 2 and secs >= budget_s:
            break
    model, ncpu = host_cpu()
    return dict(value=edges / secs / 1e9, unit="GTEPS", cores=1, kind="port",
                sample=f"{done} of the {len(srcs)} sources, serial oracle BFS (C, -O3), "
                       f"{secs:.1f} s of CPU time on one core of {model} ({ncpu} logical cores)",
                levels_bit_exact_on_sample=exact)


def other_configs(dfq):
    """The Amazon-like directed graph (source 0) and the 4096^2 grid with 2%
    of its edges dropped (source 0): one launch each after a warmup, event
    time; the grid is barrier-latency bound, so its figure is us per level."""
    out = {}
    c1 = dfq.Csr.amazon_like()
    with dfq.DeviceGraph(c1) as g:
        g.bfs(0)
        r = g.bfs(0)
        e, _, _ = c1.reached_work(r.level)
        out["amazon_like"] = dict(vertices=c1.n, csr_entries=c1.m, levels=r.depth, ms=round(r.kernel_ms, 4),
                                  gteps=round(e / (r.kernel_ms * 1e-3) / 1e9, 3))
    c3 = dfq.Csr.grid2d(4096, 4096, 0.02, 4096)
    with dfq.DeviceGraph(c3) as g:
        g.set_config(direction_opt=0)
        g.bfs(0)
        r = g.bfs(0)
        out["grid_4096"] = dict(vertices=c3.n, csr_entries=c3.m, levels=r.depth, ms=round(r.kernel_ms, 3),
                                us_per_level=round(r.kernel_ms * 1e3 / r.depth, 2))
    return out


def load_traffic(name):
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def batch_scales(count, seed, lo=10, hi=20, exponent=1.5):
    """Scales of `count` graphs: vertex counts from p(n) ~ n^-exponent on
    [2^lo, 2^hi], i.e. P(scale = s) ~ 2^(-(exponent - 1) s), seeded."""
    rng = np.random.default_rng(seed)
    s = np.arange(lo, hi + 1)
    w = 2.0 ** (-(exponent - 1.0) * (s - lo))
    return rng.choice(s, size=count, p=w / w.sum())


def run_batch(args, dfq):
    """Batched form: many graphs x sources in one persistent launch against
    the same queries issued as back-to-back plain launches."""
    t0 = time.perf_counter()
    scales = batch_scales(args.batch_graphs, seed=4)
    graphs = [dfq.Csr.rmat(int(s), 8, 0.57, 0.19, 0.19, seed=1000 + i) for i, s in enumerate(scales)]
    gi, so = [], []
    for i, g in enumerate(graphs):
        for s in g.pick_sources(args.batch_sources, seed=i):
            gi.append(i)
            so.append(int(s))
    gi, so = np.array(gi), np.array(so)
    nq = len(gi)
    log(f"{len(graphs)} graphs, {sum(g.n for g in graphs)} vertices, {sum(g.m for g in graphs)} "
        f"entries, {nq} queries, built in {time.perf_counter() - t0:.1f}s")
    with dfq.Batch(graphs) as b:
        return batch_on(args, dfq, b, graphs, gi, so)


def batch_on(args, dfq, b, graphs, gi, so):
    nq = len(gi)
    b.set_queries(gi, so)
    t0 = time.perf_counter()
    tune = b.autotune(repeats=1)
    log(f"autotune {time.perf_counter() - t0:.1f}s: block {tune['block_size']}, grid-tier team xcds "
        f"{tune['team_xcds']}, {tune['candidates']} candidates")
    for e in tune["buckets"]:
        log("  bucket", e["bucket"], "queries", e["queries"], e["plan"], f"{e['sub_batch_ms']:.3f} ms")
    for _ in range(args.warmup):
        b.run()
    ts, replays = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = b.run()
        ts.append(st.kernel_ms)
        replays.append(bool(st.graph_replay))
    wall_batch = (time.perf_counter() - t0) * 1e3 / args.steps
    # back-to-back plain launches of the same queries (one resident graph each)
    dgs = [dfq.DeviceGraph(g) for g in graphs]
    for dg in dgs:
        dg.set_config(direction_opt=1)
    dbuf = dfq.DeviceBuffer(4 * max(g.n for g in graphs))
    for q in range(min(nq, 64)):
        dgs[gi[q]].bfs_device(int(so[q]), dbuf.ptr)
    dfq.synchronize()
    t0 = time.perf_counter()
    plain_kernel = 0.0
    for q in range(nq):
        plain_kernel += dgs[gi[q]].bfs_device(int(so[q]), dbuf.ptr).kernel_ms
    plain_wall = (time.perf_counter() - t0) * 1e3
    # parity: every query's levels against the plain engine
    exact = True
    edges = 0
    for q in range(nq):
        lv = b.levels(q)
        ref = dgs[gi[q]].bfs(int(so[q])).level
        exact &= bool(np.array_equal(lv, ref))
        edges += graphs[gi[q]].reached_work(lv)[0] / 2
    for dg in dgs:
        dg.close()
    dbuf.free()
    kernel_ms = float(np.mean(ts))
    qps = nq / (kernel_ms * 1e-3)
    plain_qps = nq / (plain_wall * 1e-3)
    out = {
        "metric": "batched BFS queries/sec vs back-to-back plain launches",
        "value": round(qps, 1), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall_batch, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"{len(graphs)} undirected RMAT graphs, edgefactor 8, scales "
                               f"10..20 with p(n)~n^-1.5, {args.batch_sources} sources each",
                   "queries": nq, "vertices_total": int(sum(g.n for g in graphs)),
                   "graph_replay": all(replays), "block_size": st.block_size,
                   "grid_tier_xcds_per_team": st.team_xcds},
        "batch_kernel_ms": round(kernel_ms, 3),
        # every query of a batch completes with its launch: a query's
        # latency is the launch time, the same for all of them
        "latency_p50_ms": round(float(np.percentile(ts, 50)), 3),
        "latency_p99_ms": round(float(np.percentile(ts, 99)), 3),
        "plain_back_to_back_ms": round(plain_wall, 3),
        "plain_kernel_sum_ms": round(plain_kernel, 3),
        "plain_queries_per_s": round(plain_qps, 1),
        "speedup_vs_plain": round(qps / plain_qps, 2),
        "gteps_batch": round(edges / (kernel_ms * 1e-3) / 1e9, 2),
        "levels_bit_exact_vs_plain": exact,
        "autotune": [{"bucket": e["bucket"], "queries": e["queries"],
                      "tier": "grid" if e["plan"]["tier"] else "workgroup",
                      "medium_max": e["plan"]["medium_max"],
                      "direction_opt": e["plan"]["direction_opt"]} for e in tune["buckets"]],
    }
    return out


def run_levels(args, dfq):
    """Per-level counters of the profiling kernel (dfq_bfs_profile): every
    level of the first --level-sources searches of the headline graph in
    --mode, each next to the plain launch's time, and the 4096^2 grid's
    levels summarised by frontier size (the profiling kernel adds per-level
    atomics and timestamps, so its times run above the plain launch's).
    For the grid, each level's time splits into the slowest workgroup's
    busy time (its share of the level up to its barrier arrival) and the
    rest: the barrier from the last arrival to the release workgroup 0
    sees, and the start of the next level.  "rmat_multi": the levels of
    one multi-source traversal of the --sources sources (at most 64), the
    search that group tasks of the served form run."""
    out = {"mode": args.mode, "rmat": [], "rmat_multi": {}, "grid": {}}
    csr = dfq.Csr.rmat(args.scale, args.edgefactor, 0.57, 0.19, 0.19, seed=22)
    srcs = csr.pick_sources(args.sources, seed=2022)
    keep = ("frontier", "edges", "probes", "atomics", "claims", "items", "bottom_up", "level_ns",
            "busy_max_ns", "bin_small", "bin_medium", "bin_huge", "bu_candidates")
    with dfq.DeviceGraph(csr) as g:
        g.set_config(**MODES[args.mode])
        for s in srcs[:args.level_sources]:
            g.bfs(int(s))
            plain = g.bfs(int(s))
            rows, st = g.profile(int(s))
            out["rmat"].append({"source": int(s), "plain_ms": round(plain.kernel_ms, 4),
                                "profiled_ms": round(st.kernel_ms, 4),
                                "levels": [{k: r[k] for k in keep} for r in rows]})
        ms_src = srcs[:64]
        with dfq.DeviceBuffer(4 * csr.n * len(ms_src)) as d:
            g.bfs_device_multi(ms_src, d.ptr)
            _, ms = g.bfs_device_multi(ms_src, d.ptr)
            rows, _ = g.multi_profile(ms_src, d.ptr)
        out["rmat_multi"] = {"sources": len(ms_src), "traversal_ms": round(ms, 3), "levels": rows}
    c3 = dfq.Csr.grid2d(4096, 4096, 0.02, 4096)
    with dfq.DeviceGraph(c3) as g:
        g.set_config(direction_opt=0)
        g.bfs(0)
        plain = g.bfs(0)
        rows, st = g.profile(0, max_levels=8192)
        fr = np.array([r["frontier"] for r in rows])
        ns = np.array([r["level_ns"] for r in rows], dtype=np.float64)
        busy = np.array([r["busy_max_ns"] for r in rows], dtype=np.float64)
        out["grid"] = {"plain_ms": round(plain.kernel_ms, 3), "profiled_ms": round(st.kernel_ms, 3),
                       "levels": len(rows), "by_frontier": []}
        for lo, hi in [(0, 64), (64, 256), (256, 1024), (1024, 2048), (2048, 4096), (4096, 1 << 30)]:
            m = (fr >= lo) & (fr < hi)
            if m.any():
                out["grid"]["by_frontier"].append({"frontier": [lo, hi], "levels": int(m.sum()),
                                                   "profiled_us_mean": round(float(ns[m].mean()) / 1e3, 2),
                                                   "busy_max_us_mean": round(float(busy[m].mean()) / 1e3, 2),
                                                   "barrier_us_mean": round(float((ns[m] - busy[m]).mean()) / 1e3,
                                                                            2)})
    return out


SERVE_SCALES = (18, 19, 20, 21, 22, 22, 23, 24)


def serve_trace(graphs, scales, nq, rate_qps, seed):
    """Queries of the served mix: graph i with p ~ 1/scale_i, a uniform source
    with degree > 0 in it; Poisson arrivals at rate_qps (all at 0 if None)."""
    rng = np.random.default_rng(seed)
    p = 1.0 / np.array(scales, float)
    gi = rng.choice(len(graphs), size=nq, p=p / p.sum()).astype(np.uint32)
    so = np.zeros(nq, np.uint32)
    for g in range(len(graphs)):
        idx = np.nonzero(gi == g)[0]
        if idx.size:
            so[idx] = graphs[g].pick_sources(int(idx.size), seed=seed * 100 + g)
    arrival = np.cumsum(rng.exponential(1e6 / rate_qps, nq)) if rate_qps else np.zeros(nq)
    return gi, so, arrival


def run_serve(args, dfq):
    """Served form: resident graphs, one persistent launch of XCD teams,
    continuous batching under Poisson arrivals, against the same queries
    issued as back-to-back plain launches."""
    t0 = time.perf_counter()
    scales = SERVE_SCALES
    graphs = [dfq.Csr.rmat(s, 16, 0.57, 0.19, 0.19, seed=5000 + i) for i, s in enumerate(scales)]
    log(f"{len(graphs)} graphs, {sum(g.n for g in graphs)} vertices, {sum(g.m for g in graphs)} entries, "
        f"built in {time.perf_counter() - t0:.1f}s")

    # back-to-back plain launches on a sample of the mix
    pg, ps, _ = serve_trace(graphs, scales, args.serve_plain, None, seed=7)
    dgs = [dfq.DeviceGraph(g) for g in graphs]
    for dg in dgs:
        dg.set_config(direction_opt=1)
    dbuf = dfq.DeviceBuffer(4 * max(g.n for g in graphs))
    for i, dg in enumerate(dgs):
        dg.bfs_device(int(graphs[i].pick_sources(1, seed=1)[0]), dbuf.ptr)
    dfq.synchronize()
    plain = np.zeros((len(pg), 2), np.uint64)
    t0 = time.perf_counter()
    for q in range(len(pg)):
        st = dgs[pg[q]].bfs_device(int(ps[q]), dbuf.ptr)
        plain[q] = (st.reached, st.depth)
    plain_wall = time.perf_counter() - t0
    plain_qps = len(pg) / plain_wall
    k_full = min(24, len(pg))
    plain_levels = [dgs[pg[q]].bfs(int(ps[q])).level for q in range(k_full)]
    log(f"plain back-to-back: {len(pg)} queries in {plain_wall:.2f}s = {plain_qps:.1f} q/s")
    for dg in dgs:
        dg.close()
    dbuf.free()
    del dgs

    out = None
    for groups in ([True, False] if args.serve_groups else [False]):
        shapes = args.serve_group_shapes if groups else args.serve_shapes
        with dfq.Server(graphs, xcds_per_team=shapes[0][0], block_size=shapes[0][1], groups=groups) as srv:
            r = serve_on(args, srv, graphs, scales, pg, ps, plain, plain_qps, plain_levels, k_full, shapes, groups)
        log(f"serve {'groups' if groups else 'single'}: {r['value']} q/s, {r['speedup_vs_plain']}x, "
            f"p99 {r['poisson']['p99_ms']} ms")
        if out is None:
            out = r
        else:
            out["single_query_tasks"] = {k: r[k] for k in ("value", "speedup_vs_plain", "results_exact_vs_plain",
                                                           "gteps_served", "poisson", "shape_sweep_qps")}
            out["single_query_tasks"]["config"] = {k: r["config"][k] for k in ("xcds_per_team", "teams",
                                                                              "block_size")}
    return out


def serve_on(args, srv, graphs, scales, pg, ps, plain, plain_qps, plain_levels, k_full, shapes, groups):
    # team width: the saturated rate of a separate trace (all arrivals at
    # t = 0, up to 20,000 queries) at each shape; a short trace favours wide
    # teams, since the last traversals of the largest graph dominate it.
    # Every shape's results on the plain sample are checked against it.
    sweep = {}
    exact = True
    gs, ss, _ = serve_trace(graphs, scales, min(args.serve_queries, 20000), None, seed=17)
    for shape in shapes:
        srv.set_shape(*shape)
        srv.start()
        recs, _ = srv.replay(pg, ps, np.zeros(len(pg)))
        exact &= bool(np.array_equal(recs["reached"], plain[:, 0]) and
                      np.array_equal(recs["depth"].astype(np.uint64), plain[:, 1]))
        recs, _ = srv.replay(gs, ss, np.zeros(len(gs)))
        sweep[shape] = len(gs) / (recs["done_us"].max() * 1e-6)
        srv.stop()
        log(f"  {'groups' if groups else 'single'}, xcds/team {shape[0]}, block {shape[1]}: "
            f"{sweep[shape]:.1f} q/s saturated on {len(gs)} queries")
    best = max(sweep, key=sweep.get)
    srv.set_shape(*best)
    srv.start()
    info_s = srv.info()
    # full level arrays of a few queries, served in the mix, against the plain engine
    _, outs = srv.replay(pg[:k_full], ps[:k_full], np.zeros(k_full), keep_levels=True)
    exact &= all(np.array_equal(a, b) for a, b in zip(outs, plain_levels))

    # saturation: every arrival at t = 0, admission paces the mix
    gi, so, _ = serve_trace(graphs, scales, args.serve_queries, None, seed=11)
    recs, _ = srv.replay(gi, so, np.zeros(len(gi)))
    sat_s = recs["done_us"].max() * 1e-6
    sat_qps = len(gi) / sat_s
    sat_gteps = float(recs["edges"].sum()) / 2 / sat_s / 1e9
    log(f"saturation: {len(gi)} queries in {sat_s:.2f}s = {sat_qps:.1f} q/s, {sat_gteps:.1f} GTEPS")

    # Poisson arrivals at a fraction of the saturated rate: latency
    rate = args.serve_load * sat_qps
    gi2, so2, arr = serve_trace(graphs, scales, args.serve_queries, rate, seed=13)
    recs2, _ = srv.replay(gi2, so2, arr)
    lat = (recs2["done_us"] - recs2["arrival_us"]) / 1e3
    wait = (recs2["admit_us"] - recs2["arrival_us"]) / 1e3
    span = (recs2["done_us"].max() - recs2["arrival_us"].min()) * 1e-6
    srv.stop()
    srv.close()
    lat_by_scale = {}
    for g, sc in enumerate(scales):
        m = gi2 == g
        if m.any():
            lat_by_scale[f"g{g}_s{sc}"] = {"p50_ms": round(float(np.percentile(lat[m], 50)), 3),
                                           "p99_ms": round(float(np.percentile(lat[m], 99)), 3),
                                           "service_ms_mean": round(float(recs2["service_us"][m].mean()) / 1e3, 3)}
    out = {
        "metric": "served BFS queries/sec vs back-to-back plain launches, with p50/p99 latency",
        "tasks": "group tasks: queries of one graph waiting together share one bit-parallel "
                 "multi-source traversal (up to 64)" if groups else "one search per team task",
        "value": round(sat_qps, 1), "unit": "queries/s", "n_gpus": 1, "steps": 1,
        "warmup": len(args.serve_shapes), "ms_per_step": round(sat_s * 1e3, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": "8 resident undirected RMAT graphs (scales 18,19,20,21,22,22,23,24, "
                               "edgefactor 16); graph i drawn with p ~ 1/scale_i, uniform source with "
                               "degree > 0",
                   "queries": len(gi), "vertices_total": int(sum(g.n for g in graphs)),
                   "csr_entries_total": int(sum(g.m for g in graphs)),
                   "xcds_per_team": best[0], "teams": info_s["teams"], "block_size": info_s["block_size"],
                   "arena_bytes": info_s["arena_bytes"]},
        "plain_queries_per_s": round(plain_qps, 1),
        "plain_sample": len(pg),
        "speedup_vs_plain": round(sat_qps / plain_qps, 2),
        "gteps_served": round(sat_gteps, 1),
        "poisson": {"rate_qps": round(rate, 1), "load": args.serve_load, "queries": len(gi2),
                    "achieved_qps": round(len(gi2) / span, 1),
                    "p50_ms": round(float(np.percentile(lat, 50)), 3),
                    "p99_ms": round(float(np.percentile(lat, 99)), 3),
                    "p999_ms": round(float(np.percentile(lat, 99.9)), 3),
                    "admission_wait_p99_ms": round(float(np.percentile(wait, 99)), 3),
                    "by_graph": lat_by_scale},
        "shape_sweep_qps": {f"{k[0]}x{k[1]}": round(v, 1) for k, v in sweep.items()},
        "results_exact_vs_plain": exact,
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edgefactor", type=int, default=16)
    ap.add_argument("--sources", type=int, default=64)
    ap.add_argument("--mode", choices=sorted(MODES), default="hybrid")
    ap.add_argument("--also", choices=sorted(MODES) + ["none"], default="topdown",
                    help="second mode reported beside the headline")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the Amazon-like and grid lines")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the batched and served objects of the headline line")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the library's streaming copy of BYTES (counter calibration)")
    ap.add_argument("--workload", choices=["plain", "batch", "serve", "levels"], default="plain",
                    help="plain: the RMAT-22 headline; batch: batched queries (one launch); "
                         "serve: resident graphs under Poisson arrivals; levels: per-level "
                         "counters (dfq_bfs_profile) of --level-sources RMAT sources in --mode "
                         "and of the 4096^2 grid")
    ap.add_argument("--level-sources", type=int, default=2)
    ap.add_argument("--batch-graphs", type=int, default=512)
    ap.add_argument("--batch-sources", type=int, default=4)
    ap.add_argument("--serve-queries", type=int, default=100000)
    ap.add_argument("--serve-plain", type=int, default=2000, help="plain back-to-back sample")
    ap.add_argument("--serve-load", type=float, default=0.8,
                    help="Poisson rate as a fraction of the saturated rate")
    ap.add_argument("--serve-shapes", type=int, nargs="+", default=[1, 2, 4],
                    help="XCDs per team to compare (the fastest serves the trace)")
    ap.add_argument("--serve-group-shapes", type=str, nargs="+", default=["1x512", "1x1024", "2x1024", "8x1024"],
                    help="XCDs-per-team x workgroup size for group-mode serving")
    ap.add_argument("--no-serve-groups", dest="serve_groups", action="store_false",
                    help="serve single-query tasks only")
    ap.add_argument("--then-copy", type=int, default=0, metavar="BYTES",
                    help="after the benchmark, run the streaming copy of BYTES once (so one "
                         "counter pass covers the BFS and its calibration)")
    args = ap.parse_args()
    args.serve_shapes = [(x, 1024) for x in args.serve_shapes]
    args.serve_group_shapes = [tuple(int(v) for v in x.split("x")) for x in args.serve_group_shapes]
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build runs the search on one "
              f"GPU (multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import dfqbfs as dfq
    info = dfq.device_info()
    log("device", info)
    if args.copy_calibration:
        src, dst = dfq.DeviceBuffer(args.copy_calibration), dfq.DeviceBuffer(args.copy_calibration)
        ms = [dfq.stream_copy(dst, src, args.copy_calibration) for _ in range(3)]
        print(json.dumps({"copy_bytes": args.copy_calibration, "launches": 3,
                          "ms": [round(x, 4) for x in ms],
                          "GBps_read_plus_write": round(2 * args.copy_calibration / min(ms) / 1e6, 1)}))
        return
    if args.workload == "batch":
        print(json.dumps(run_batch(args, dfq)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(run_serve(args, dfq)), flush=True)
        return
    if args.workload == "levels":
        print(json.dumps(run_levels(args, dfq)), flush=True)
        return
    t0 = time.perf_counter()
    csr = dfq.Csr.rmat(args.scale, args.edgefactor, 0.57, 0.19, 0.19, seed=22)
    log(f"graph n={csr.n} m={csr.m} built in {time.perf_counter() - t0:.1f}s")
    srcs = csr.pick_sources(args.sources, seed=2022)
    ns = len(srcs)
    out = {}
    with dfq.DeviceGraph(csr) as g, dfq.DeviceBuffer(4 * csr.n * ns) as d_level:
        g.set_config(**MODES[args.mode])
        # Untimed: the same sequence launch with parents, every search's
        # parent tree certified (levels + tree rules over every vertex/edge).
        with dfq.DeviceBuffer(4 * csr.n * ns) as d_parent:
            g.bfs_device_seq(srcs, d_level.ptr, level_stride=csr.n, d_parent=d_parent.ptr,
                             parent_stride=csr.n)
            for i, s in enumerate(srcs):
                lv = d_level.to_numpy(count=csr.n, offset=i * csr.n)
                pa = d_parent.to_numpy(count=csr.n, offset=i * csr.n)
                rep = csr.validate(int(s), lv, pa)
                if rep["errors"]:
                    raise SystemExit(f"BFS certificate failed for source {s}: {rep}")
        log("certificate (levels + parent tree) ok on all sources")

        res = run_mode(dfq, g, srcs, None, args.steps, args.warmup, d_level)
        # the levels of the last timed launch: certificate and reached work
        # of every search, and (cpu_baseline) bit-exactness on a sample
        levels = [d_level.to_numpy(count=csr.n, offset=i * csr.n) for i in range(ns)]
        work = []
        for s, lv in zip(srcs, levels):
            rep = csr.validate(int(s), lv)
            if rep["errors"]:
                raise SystemExit(f"timed launch: level certificate failed for source {s}: {rep}")
            work.append(csr.reached_work(lv))
        res = price(res, work, csr.n)
        log(args.mode, res)
        each = price(run_mode(dfq, g, srcs, None, args.steps, args.warmup, d_level, launch="each"),
                     work, csr.n)
        log(args.mode, "one launch per search", each)
        extra = {"one_launch_per_search": dict(gteps_hmean=round(each["gteps_hmean"], 3),
                                               ms_per_bfs_mean=round(each["ms_per_bfs_mean"], 4),
                                               roofline_frac=round(each["achieved_gbs"] / HBM_PEAK_GBS, 4))}
        if args.also != "none" and args.also != args.mode:
            g.set_config(**MODES[args.also])
            other = price(run_mode(dfq, g, srcs, None, args.steps, args.warmup, d_level), work, csr.n)
            log(args.also, other)
            extra[args.also] = dict(gteps_hmean=round(other["gteps_hmean"], 3),
                                    ms_per_bfs_mean=round(other["ms_per_bfs_mean"], 4),
                                    roofline_frac=round(other["achieved_gbs"] / HBM_PEAK_GBS, 4))

    traffic = load_traffic(f"traffic_rmat{args.scale}_{args.mode}.json")
    out.update({
        "metric": "GTEPS + % HBM roofline (RMAT-22)",
        "value": round(res["gteps_hmean"], 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["wall_ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"RMAT scale {args.scale}, edgefactor {args.edgefactor}, "
                               f"a/b/c=0.57/0.19/0.19, undirected, {args.sources} sources/step "
                               f"in one launch, levels output",
                   "vertices": csr.n, "csr_entries": csr.m, "mode": args.mode,
                   "kernel_config": MODES[args.mode]},
        "roofline": {"bound": "hbm", "achieved": round(res["achieved_gbs"], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(res["achieved_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic["bytes_per_launch"] if traffic else None},
        "gteps_median": round(res["gteps_median"], 3),
        "ms_per_bfs_mean": round(res["ms_per_bfs_mean"], 4),
        "b_alg_bytes_mean": int(res["b_alg_mean"]),
        "b_alg_bytes_per_launch": int(res["b_alg_step"]),
        "launch_ms_mean": round(res["kernel_ms_total"], 4),
        "timed_output_checked": "levels of the last timed launch (dfq_bfs_seq_kernel, stride n): "
                                "certificate on every search; parents of an untimed launch of the "
                                "same kernel and config certified on every search",
    })
    out.update(extra)
    if not args.no_extra:
        out["other_configs"] = other_configs(dfq)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(csr, srcs, levels, args.cpu_budget)
    del levels
    if not args.no_forms:
        b = run_batch(args, dfq)
        out["batch"] = {k: b[k] for k in ("value", "unit", "plain_queries_per_s", "speedup_vs_plain",
                                          "levels_bit_exact_vs_plain", "latency_p50_ms",
                                          "latency_p99_ms", "gteps_batch")}
        out["batch"]["config"] = b["config"]
        sv = run_serve(args, dfq)
        out["serve"] = {k: sv[k] for k in ("value", "unit", "plain_queries_per_s", "speedup_vs_plain",
                                           "results_exact_vs_plain", "gteps_served", "tasks")}
        if "single_query_tasks" in sv:
            out["serve"]["single_query_tasks"] = {k: sv["single_query_tasks"][k]
                                                  for k in ("value", "speedup_vs_plain", "results_exact_vs_plain")}
        out["serve"]["poisson"] = {k: sv["poisson"][k] for k in ("rate_qps", "queries", "achieved_qps",
                                                                 "p50_ms", "p99_ms")}
        out["serve"]["config"] = sv["config"]
    print(json.dumps(out), flush=True)
    if args.then_copy:
        a, b = dfq.DeviceBuffer(args.then_copy), dfq.DeviceBuffer(args.then_copy)
        log("copy", args.then_copy, "bytes:", round(dfq.stream_copy(b, a, args.then_copy), 4), "ms")


if __name__ == "__main__":
    main()
