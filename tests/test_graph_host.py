"""Host-side library (include/dfq_graph.h): CSR construction, file formats,
generators, source selection, BFS certificate validator.  CPU only."""
import os

import numpy as np
import pytest

from conftest import UNREACHED, csr_from_edges_np, load_golden


def test_csr_from_edges_flags(dfq):
    src = [0, 0, 1, 2, 2, 3]
    dst = [1, 1, 2, 2, 0, 0]
    g = dfq.Csr.from_edges(src, dst, 4)
    np.testing.assert_array_equal(g.row_offsets, [0, 2, 3, 5, 6])
    np.testing.assert_array_equal(g.col_idx, [1, 1, 2, 0, 2, 0])
    g = dfq.Csr.from_edges(src, dst, 4, dfq.CSR_DEDUP | dfq.CSR_DROP_SELF)
    np.testing.assert_array_equal(g.row_offsets, [0, 1, 2, 3, 4])
    np.testing.assert_array_equal(g.col_idx, [1, 2, 0, 0])
    g = dfq.Csr.from_edges(src, dst, 4, dfq.CSR_SYMMETRIZE | dfq.CSR_DEDUP | dfq.CSR_DROP_SELF)
    # undirected edges {0,1},{1,2},{0,2},{0,3}
    np.testing.assert_array_equal(g.row_offsets, [0, 3, 5, 7, 8])
    np.testing.assert_array_equal(g.col_idx, [1, 2, 3, 0, 2, 0, 1, 0])
    with pytest.raises(dfq.DfqError) as e:
        dfq.Csr.from_edges([0, 5], [1, 1], 4)
    assert e.value.code == dfq.ERR_RANGE


def test_permutation_is_isomorphism(dfq):
    rng = np.random.default_rng(3)
    src, dst = rng.integers(0, 500, 3000), rng.integers(0, 500, 3000)
    a = dfq.Csr.from_edges(src, dst, 500, dfq.CSR_UNDIRECTED & ~dfq.CSR_PERMUTE)
    b = dfq.Csr.from_edges(src, dst, 500, dfq.CSR_UNDIRECTED, perm_seed=9)
    assert a.m == b.m
    assert sorted(a.degrees()) == sorted(b.degrees())
    assert not np.array_equal(a.degrees(), b.degrees())


def test_rmat_deterministic_and_undirected(dfq):
    a = dfq.Csr.rmat(12, 8, seed=5)
    b = dfq.Csr.rmat(12, 8, seed=5)
    c = dfq.Csr.rmat(12, 8, seed=6)
    np.testing.assert_array_equal(a.col_idx, b.col_idx)
    np.testing.assert_array_equal(a.row_offsets, b.row_offsets)
    assert a.m != c.m or not np.array_equal(a.col_idx, c.col_idx)
    ro, ci = a.row_offsets, a.col_idx
    rows = np.repeat(np.arange(a.n), np.diff(ro))
    fwd = set(zip(rows.tolist(), ci.tolist()))
    assert all((v, u) in fwd for u, v in fwd)           # symmetric
    assert all(u != v for u, v in fwd)                   # no loops
    assert len(fwd) == a.m                               # no duplicates
    # skewed degrees: a = .57 puts mass on low ids before permutation
    assert a.degrees().max() > 20 * a.degrees().mean()


def test_rmat_scale22_sizes(dfq):
    """Config 2 graph: sizes close to the survey's seed-1 measurement."""
    g = dfq.Csr.rmat(22, 16, seed=22)
    deg = g.degrees()
    assert g.n == 1 << 22
    assert 127_500_000 < g.m < 129_000_000
    assert 2_380_000 < (deg > 0).sum() < 2_410_000
    assert 200 < (deg > 16384).sum() < 320


def test_amazon_like_rules(dfq):
    g = dfq.Csr.amazon_like(20000, seed=11)
    deg = g.degrees()
    assert deg.min() >= 1 and deg.max() <= 5
    ro, ci = g.row_offsets, g.col_idx
    rows = np.repeat(np.arange(g.n), deg)
    assert (ci != rows).all()
    for v in range(0, g.n, 997):
        r = ci[ro[v]:ro[v + 1]]
        assert len(set(r.tolist())) == len(r)
    near = np.abs(ci.astype(np.int64) - rows) <= 2000
    assert 0.65 < near.mean() < 0.8  # 70% local + uniform draws that land nearby
    h = dfq.Csr.amazon_like(20000, seed=11)
    np.testing.assert_array_equal(h.col_idx, g.col_idx)


def test_grid2d(dfq):
    g = dfq.Csr.grid2d(30, 20, 0.0)
    assert g.n == 600 and g.m == 2 * (29 * 20 + 30 * 19)
    h = dfq.Csr.grid2d(300, 200, 0.02, seed=1)
    full = 2 * (299 * 200 + 300 * 199)
    assert 0.97 * full < h.m < 0.99 * full
    ro, ci = h.row_offsets, h.col_idx
    rows = np.repeat(np.arange(h.n), np.diff(ro))
    fwd = set(zip(rows.tolist(), ci.tolist()))
    assert all((v, u) in fwd for u, v in fwd)


def test_snap_roundtrip_and_header(dfq, tmp_path):
    g = dfq.Csr.amazon_like(500, seed=2)
    p = str(tmp_path / "g.txt")
    g.save_snap(p)
    head = open(p).read().splitlines()[:2]
    assert head[0] == f"# Nodes: {g.n} Edges: {g.m}"
    assert head[1].startswith("# FromNodeId")
    h = dfq.Csr.load_snap(p)
    np.testing.assert_array_equal(h.row_offsets, g.row_offsets)
    np.testing.assert_array_equal(h.col_idx, g.col_idx)


def test_snap_parsing_edge_cases(dfq, tmp_path):
    p = tmp_path / "x.txt"
    p.write_text("# Directed graph\n# Nodes: 6 Edges: 4\n# FromNodeId\tToNodeId\n"
                 "0\t1\n1 2\r\n\n  4\t5\t7\n5\t0\n")
    g = dfq.Csr.load_snap(str(p))
    assert g.n == 6 and g.m == 4
    np.testing.assert_array_equal(g.col_idx, [1, 2, 5, 0])
    p.write_text("0 1\n3 2\n")  # no header: n = max id + 1
    assert dfq.Csr.load_snap(str(p)).n == 4
    p.write_text("# Nodes: 3 Edges: 1\n0 3\n")
    with pytest.raises(dfq.DfqError) as e:
        dfq.Csr.load_snap(str(p))
    assert e.value.code == dfq.ERR_RANGE
    p.write_text("# Nodes: 3 Edges: 1\n0 x\n")
    with pytest.raises(dfq.DfqError) as e:
        dfq.Csr.load_snap(str(p))
    assert e.value.code == dfq.ERR_FORMAT
    with pytest.raises(dfq.DfqError) as e:
        dfq.Csr.load_snap(str(tmp_path / "missing.txt"))
    assert e.value.code == dfq.ERR_IO


def test_dat_roundtrip_and_reference_layout(dfq, tmp_path):
    g = dfq.Csr.rmat(8, 4, seed=3)
    p = str(tmp_path / "g.dat")
    g.save_dat(p, source=7)
    lines = open(p).read().split("\n")
    assert lines[0] == f"{g.n} {g.m} 7"
    h, s = dfq.Csr.load_dat(p)
    assert s == 7
    np.testing.assert_array_equal(h.row_offsets, g.row_offsets)
    np.testing.assert_array_equal(h.col_idx, g.col_idx)
    # reference rows may be listed with arbitrary starts (Node{start,length})
    q = tmp_path / "r.dat"
    q.write_text("3 4 0\n2 2\n0 1\n1 1\n2 5\n1 5\n0 5\n1 5\n")
    h, s = dfq.Csr.load_dat(str(q))
    np.testing.assert_array_equal(h.row_offsets, [0, 2, 3, 4])
    np.testing.assert_array_equal(h.col_idx, [0, 1, 2, 1])
    q.write_text("3 4 5\n0 1\n1 1\n2 2\n1 1\n2 1\n0 1\n1 1\n")  # source >= N
    with pytest.raises(dfq.DfqError) as e:
        dfq.Csr.load_dat(str(q))
    assert e.value.code == dfq.ERR_RANGE


def test_csr_bin_roundtrip(dfq, tmp_path):
    g = dfq.Csr.rmat(10, 8, seed=4)
    p = str(tmp_path / "g.csr")
    g.save_csr_bin(p)
    h = dfq.Csr.load_csr_bin(p)
    np.testing.assert_array_equal(h.col_idx, g.col_idx)
    open(p, "r+b").write(b"BADMAGIC")
    with pytest.raises(dfq.DfqError):
        dfq.Csr.load_csr_bin(p)


def test_pick_sources(dfq):
    g = dfq.Csr.rmat(12, 4, seed=8)
    s = g.pick_sources(64, seed=2022)
    assert len(set(s.tolist())) == 64
    assert (g.degrees()[s] > 0).all()
    np.testing.assert_array_equal(s, g.pick_sources(64, seed=2022))


def test_validator_accepts_oracle_and_rejects_planted_errors(dfq, oracle):
    gd = load_golden("rmat12_ef8")
    g = dfq.Csr.from_arrays(gd["row_offsets"], gd["col_idx"])
    s = int(gd["sources"][0])
    lv, par = oracle.bfs(gd["row_offsets"], gd["col_idx"], s, parents=True)
    assert g.validate(s, lv, par)["errors"] == 0
    assert g.validate(s, lv)["errors"] == 0
    reached = np.flatnonzero((lv != UNREACHED) & (lv > 0))
    v = int(reached[len(reached) // 2])
    bad = lv.copy()
    bad[v] += 1                                 # one level too deep
    assert g.validate(s, bad, par)["errors"] > 0
    bad = lv.copy()
    bad[v] = UNREACHED                          # dropped vertex
    r = g.validate(s, bad)
    assert r["missed_vertex"] > 0
    badp = par.copy()
    badp[v] = v                                 # parent not an edge / wrong level
    assert g.validate(s, lv, badp)["bad_parent"] == 1
    deep = lv.copy()
    deep[lv == lv.max()] -= 1                   # levels too shallow: no witness
    assert g.validate(s, deep)["errors"] > 0


def test_reached_work(dfq, oracle):
    g = dfq.Csr.rmat(11, 8, seed=1)
    s = int(g.pick_sources(1, 1)[0])
    lv = oracle.bfs(g.row_offsets, g.col_idx, s)
    e, v, d = g.reached_work(lv)
    mask = lv != UNREACHED
    assert v == mask.sum() and e == g.degrees()[mask].sum() and d == lv[mask].max() + 1


def test_library_rejects_bad_arguments(dfq):
    with pytest.raises(dfq.DfqError):
        dfq.Csr.rmat(0, 16)
    with pytest.raises(dfq.DfqError):
        dfq.Csr.rmat(10, 16, a=0.9, b=0.2, c=0.1)
    with pytest.raises(dfq.DfqError):
        dfq.Csr.grid2d(0, 5)
    with pytest.raises(dfq.DfqError):
        dfq.Csr.from_arrays([0, 2, 1], [0, 1])


def test_crafted_files_are_refused(dfq, tmp_path):
    """Inputs that once read or wrote past a heap buffer, or threw out of the
    C ABI, now come back as status codes."""
    q = tmp_path / "c.dat"
    q.write_text("1 1 0\n18446744073709551615 2\n0 1\n")   # start + length wraps
    with pytest.raises(dfq.DfqError) as e:
        dfq.Csr.load_dat(str(q))
    assert e.value.code == dfq.ERR_RANGE
    q.write_text("2 1000000000000000000 0\n0 0\n0 0\n")    # header larger than the file
    with pytest.raises(dfq.DfqError) as e:
        dfq.Csr.load_dat(str(q))
    assert e.value.code in (dfq.ERR_RANGE, dfq.ERR_FORMAT)
    q.write_text("1 1 0\n0 123456789012345678901234567890\n0 1\n")  # > 64-bit number
    with pytest.raises(dfq.DfqError) as e:
        dfq.Csr.load_dat(str(q))
    assert e.value.code == dfq.ERR_FORMAT
    b = tmp_path / "c.csr"
    b.write_bytes(b"DFQCSR01" + np.uint32(1).tobytes() + np.uint64(1 << 62).tobytes() + bytes(16))
    with pytest.raises(dfq.DfqError) as e:                # entry count wraps the size
        dfq.Csr.load_csr_bin(str(b))
    assert e.value.code == dfq.ERR_FORMAT
    g = dfq.Csr.rmat(8, 4, seed=3)
    g.save_csr_bin(str(b))
    data = b.read_bytes()
    b.write_bytes(data[:-4])                              # truncated body
    with pytest.raises(dfq.DfqError):
        dfq.Csr.load_csr_bin(str(b))
    s = tmp_path / "c.txt"
    s.write_text("# Nodes: 99999999999999999999999 Edges: 1\n0 1\n")
    with pytest.raises(dfq.DfqError) as e:
        dfq.Csr.load_snap(str(s))
    assert e.value.code == dfq.ERR_FORMAT


def test_host_library_under_sanitizers():
    """`make asan`: the host library built with ASan + UBSan (g++, no device
    code) runs crafted and seeded-mutated files of every format, the
    generators and the certificate check; any sanitizer report fails it."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not installed")
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dfq-bfs_amd")
    r = subprocess.run(["make", "-s", "-C", pkg, "asan"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "all checks passed" in r.stdout
