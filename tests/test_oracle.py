"""Pins the oracle (C restatement of the reference's serial BFS,
design/shared-q/atomic-sa/our-sep/upper-bound/bfs.cu:216-253) against checks
that do not come from the oracle itself: scipy-computed golden vectors,
closed forms, and Floyd-Warshall brute force."""
import itertools

import numpy as np
import pytest

from conftest import UNREACHED, csr_from_edges_np, golden_names, load_golden


@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_golden(oracle, name):
    g = load_golden(name)
    for s, expect in zip(g["sources"], g["levels"]):
        lv = oracle.bfs(g["row_offsets"], g["col_idx"], int(s))
        np.testing.assert_array_equal(lv, expect, err_msg=f"{name} source {s}")


def test_path_graph(oracle):
    n = 100
    ro, ci = csr_from_edges_np(n, np.arange(n - 1), np.arange(1, n))
    np.testing.assert_array_equal(oracle.bfs(ro, ci, 0), np.arange(n, dtype=np.uint32))
    lv = oracle.bfs(ro, ci, 50)  # directed: nothing behind the source
    assert (lv[:50] == UNREACHED).all() and (lv[50:] == np.arange(50)).all()


def test_complete_binary_tree(oracle):
    n = 1023
    child = np.arange(1, n)
    ro, ci = csr_from_edges_np(n, (child - 1) // 2, child)
    expect = np.floor(np.log2(np.arange(n) + 1)).astype(np.uint32)
    np.testing.assert_array_equal(oracle.bfs(ro, ci, 0), expect)


def test_star_and_complete(oracle):
    n = 50
    ro, ci = csr_from_edges_np(n, np.zeros(n - 1, int), np.arange(1, n))
    lv = oracle.bfs(ro, ci, 0)
    assert lv[0] == 0 and (lv[1:] == 1).all()
    src, dst = zip(*[(i, j) for i in range(n) for j in range(n) if i != j])
    ro, ci = csr_from_edges_np(n, src, dst)
    lv = oracle.bfs(ro, ci, 7)
    assert lv[7] == 0 and (np.delete(lv, 7) == 1).all()


def test_intact_grid_is_manhattan(oracle):
    W, H = 37, 23
    e = [(y * W + x, y * W + x + 1) for y in range(H) for x in range(W - 1)]
    e += [(y * W + x, (y + 1) * W + x) for y in range(H - 1) for x in range(W)]
    e = np.array(e)
    ro, ci = csr_from_edges_np(W * H, np.r_[e[:, 0], e[:, 1]], np.r_[e[:, 1], e[:, 0]])
    lv = oracle.bfs(ro, ci, 0)
    x, y = np.meshgrid(np.arange(W), np.arange(H))
    np.testing.assert_array_equal(lv, (x + y).ravel().astype(np.uint32))


def test_disconnected_self_loops_and_zero_degree_source(oracle):
    n = 10
    src = [0, 0, 1, 1, 2, 5, 6]
    dst = [1, 1, 1, 2, 0, 6, 5]  # duplicate 0->1, loop 1->1
    ro, ci = csr_from_edges_np(n, src, dst)
    lv = oracle.bfs(ro, ci, 0)
    np.testing.assert_array_equal(lv[:3], [0, 1, 2])
    assert (lv[3:] == UNREACHED).all()
    lv = oracle.bfs(ro, ci, 9)  # no out-edges
    assert lv[9] == 0 and (np.delete(lv, 9) == UNREACHED).all()


def floyd_warshall_levels(n, src, dst):
    inf = np.iinfo(np.int64).max // 4
    d = np.full((n, n), inf, np.int64)
    np.fill_diagonal(d, 0)
    for u, v in zip(src, dst):
        if u != v:
            d[u, v] = 1
    for k in range(n):
        d = np.minimum(d, d[:, k:k + 1] + d[k:k + 1, :])
    out = np.where(d >= inf, UNREACHED, d).astype(np.uint32)
    return out


@pytest.mark.parametrize("seed", range(6))
def test_brute_force_all_sources(oracle, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 40))
    m = int(rng.integers(0, 3 * n))
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    ro, ci = csr_from_edges_np(n, src, dst)
    fw = floyd_warshall_levels(n, src, dst)
    for s in range(n):
        np.testing.assert_array_equal(oracle.bfs(ro, ci, s), fw[s])


def test_parent_tree_of_oracle_is_valid(oracle):
    g = load_golden("rmat12_ef8")
    s = int(g["sources"][0])
    lv, par = oracle.bfs(g["row_offsets"], g["col_idx"], s, parents=True)
    ro, ci = g["row_offsets"], g["col_idx"]
    assert par[s] == s
    for v in np.flatnonzero(lv != UNREACHED):
        if v == s:
            continue
        p = par[v]
        assert lv[p] + 1 == lv[v]
        assert v in ci[ro[p]:ro[p + 1]]


def test_compare_counts_like_reference(oracle):
    cpu = np.array([0, 1, 2, UNREACHED, 3], np.uint32)
    gpu = np.array([0, 1, 5, 7, 3], np.uint32)
    dist, touch, err = oracle.compare(cpu, gpu)
    assert (dist, touch, err) == (3, 4, 1)


def test_planted_mistake_is_caught(oracle):
    """A plausible oracle bug (level written as parent's level, no +1) must
    fail against the golden vectors; restated here on the CPU in numpy."""
    g = load_golden("rmat12_ef8")
    s = int(g["sources"][0])
    ro, ci = g["row_offsets"], g["col_idx"]
    n = ro.size - 1
    lv = np.full(n, UNREACHED, np.uint32)
    lv[s] = 0
    frontier = [s]
    while frontier:
        nxt = []
        for u in reversed(frontier):
            for w in ci[ro[u]:ro[u + 1]]:
                if lv[w] == UNREACHED:
                    lv[w] = lv[u]  # planted: missing + 1
                    nxt.append(w)
        frontier = nxt
    assert not np.array_equal(lv, g["levels"][0])
    assert not itertools.starmap(np.array_equal, [(lv, oracle.bfs(ro, ci, s))]).__next__()
