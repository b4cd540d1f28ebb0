"""Writes the golden BFS fixtures under tests/golden/.

Each fixture holds a small graph (CSR arrays built here with numpy), a few
sources, and the expected hop distances computed by scipy.sparse.csgraph
(a third-party BFS; the reference repository ships no test vectors and no
dataset for this path).  Graph construction uses numpy's PCG64 with fixed
seeds and is independent of the library under test.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import shortest_path

HERE = os.path.dirname(os.path.abspath(__file__))
UNREACHED = np.uint32(0xFFFFFFFF)


def to_csr(n, src, dst):
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    ro = np.zeros(n + 1, np.int64)
    np.add.at(ro, src + 1, 1)
    return np.cumsum(ro), dst.astype(np.uint32)


def levels_scipy(n, ro, ci, sources):
    data = np.ones(ci.size, np.float64)
    a = csr_matrix((data, ci.astype(np.int64), ro), shape=(n, n))
    a.sum_duplicates()
    d = shortest_path(a, method="D", directed=True, unweighted=True, indices=sources)
    out = np.full((len(sources), n), UNREACHED, np.uint32)
    fin = np.isfinite(d)
    out[fin] = d[fin].astype(np.uint32)
    return out


def rmat_edges(rng, scale, ef, a=0.57, b=0.19, c=0.19):
    ne = ef << scale
    u = np.zeros(ne, np.int64)
    v = np.zeros(ne, np.int64)
    for _ in range(scale):
        r = rng.random(ne)
        bu = (r >= a + b).astype(np.int64)
        bv = (((r >= a) & (r < a + b)) | (r >= a + b + c)).astype(np.int64)
        u = (u << 1) | bu
        v = (v << 1) | bv
    return u, v


def undirected(n, u, v, rng=None, permute=True):
    keep = u != v
    u, v = u[keep], v[keep]
    a = np.concatenate([u, v])
    b = np.concatenate([v, u])
    if permute:
        p = rng.permutation(n)
        a, b = p[a], p[b]
    key = np.unique(a * n + b)
    return key // n, key % n


def write(name, n, ro, ci, sources, note):
    lv = levels_scipy(n, ro, ci, sources)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), n=np.int64(n), row_offsets=ro,
                        col_idx=ci, sources=np.asarray(sources, np.uint32), levels=lv,
                        note=np.array(note))
    reached = (lv != UNREACHED).sum(axis=1)
    print(f"{name}: n={n} m={ci.size} sources={list(sources)} reached={list(reached)} "
          f"depth={[int(x[x != UNREACHED].max()) for x in lv]}")


def main():
    # 1. directed random graph with self loops and repeated edges kept
    rng = np.random.default_rng(1)
    n = 3000
    src = rng.integers(0, n, 12000)
    dst = rng.integers(0, n, 12000)
    src = np.concatenate([src, [5, 5, 7]])
    dst = np.concatenate([dst, [5, 9, 9]])
    ro, ci = to_csr(n, src, dst)
    write("er_directed_3000", n, ro, ci, [0, 5, 17, 2999], "directed G(n,m) with loops/duplicates")

    # 2. undirected RMAT scale 12, edgefactor 8, permuted ids
    rng = np.random.default_rng(12)
    u, v = rmat_edges(rng, 12, 8)
    a, b = undirected(1 << 12, u, v, rng)
    ro, ci = to_csr(1 << 12, a, b)
    deg = np.diff(ro)
    srcs = rng.choice(np.flatnonzero(deg > 0), 4, replace=False)
    write("rmat12_ef8", 1 << 12, ro, ci, list(srcs), "undirected RMAT scale 12 ef 8 a/b/c .57/.19/.19")

    # 3. 64x64 grid, 5% of undirected edges removed
    rng = np.random.default_rng(64)
    W = 64
    e = []
    for y in range(W):
        for x in range(W):
            vtx = y * W + x
            if x + 1 < W:
                e.append((vtx, vtx + 1))
            if y + 1 < W:
                e.append((vtx, vtx + W))
    e = np.array(e)
    e = e[rng.random(len(e)) >= 0.05]
    a = np.concatenate([e[:, 0], e[:, 1]])
    b = np.concatenate([e[:, 1], e[:, 0]])
    ro, ci = to_csr(W * W, a, b)
    write("grid64_drop5", W * W, ro, ci, [0, W * W - 1, 2080], "64x64 grid, 5% edges dropped")

    # 4. amazon-like directed graph, n = 8192
    rng = np.random.default_rng(0xD0F1)
    n = 8192
    src, dst = [], []
    for vtx in range(n):
        d = int(rng.integers(1, 6))
        chosen = set()
        while len(chosen) < d:
            if rng.random() < 0.7:
                lo, hi = max(0, vtx - 2000), min(n - 1, vtx + 2000)
                t = int(rng.integers(lo, hi + 1))
            else:
                t = int(rng.integers(0, n))
            if t != vtx:
                chosen.add(t)
        for t in sorted(chosen):
            src.append(vtx)
            dst.append(t)
    ro, ci = to_csr(n, src, dst)
    write("amazon_like_8192", n, ro, ci, [0, 4096, 8191], "directed, out-degree U{1..5}, 70% local window")

    # 5. hubs at the degree-bin boundaries plus a sparse random background
    rng = np.random.default_rng(77)
    hub_degs = [1, 8, 9, 16, 17, 32, 33, 64, 65, 1024, 1025, 2048, 2049, 8192, 8193, 16384,
                16385, 32768, 32769, 40000]
    n = 50000
    src, dst = [], []
    for i, d in enumerate(hub_degs):
        hub = i * 97
        nb = rng.choice(n, d + 1, replace=False)
        nb = nb[nb != hub][:d]
        src.extend([hub] * len(nb))
        dst.extend(nb.tolist())
        # chain the hubs so one source reaches all of them
        nxt = (i + 1) * 97 if i + 1 < len(hub_degs) else 0
        src.append(hub)
        dst.append(nxt)
    bu = rng.integers(0, n, 30000)
    bv = rng.integers(0, n, 30000)
    src.extend(bu.tolist())
    dst.extend(bv.tolist())
    src, dst = np.array(src), np.array(dst)
    key = np.unique(src * n + dst)
    ro, ci = to_csr(n, key // n, key % n)
    write("bin_hubs_50000", n, ro, ci, [0, 97 * 10, 12345], "hubs of degree 8/9/.../40000 at bin edges")


if __name__ == "__main__":
    main()
