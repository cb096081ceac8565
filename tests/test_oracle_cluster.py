"""Pinning of the oracle's Louvain (src/clustering.jl, src/modularity.jl):
the karate fixture of test/test_cluster.jl, the published optimum of the
karate graph's modularity, a dense-matrix modularity, brute force over all
partitions of small graphs, and the NaN rule of group_singletons!."""
import itertools
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))


def karate():
    d = json.load(open(os.path.join(HERE, "golden", "karate.json")))
    n = d["n"]
    w = np.ones(len(d["indices"]))
    return n, np.array(d["indptr"]), np.array(d["indices"]), w, d["expected_nclusters"]


def dense_modularity(A, lab, res):
    """Q = sum_ij (A_ij - res k_i k_j / 2m) [c_i = c_j] / 2m with the
    diagonal removed (Network(snn) skips self-links)."""
    A = A.copy()
    np.fill_diagonal(A, 0.0)
    k = A.sum(0)
    m2 = A.sum()
    same = lab[:, None] == lab[None, :]
    return float(((A - res * np.outer(k, k) / m2) * same).sum() / m2)


def random_graph(n, p, seed, self_links=True):
    rng = np.random.default_rng(seed)
    U = np.triu((rng.random((n, n)) < p) * rng.uniform(0.1, 1.0, (n, n)), 1)
    A = U + U.T
    if self_links:
        np.fill_diagonal(A, 1.0)
    return A


def csc(A):
    M = sp.csc_matrix(A)
    M.sort_indices()
    return M


def test_karate_three_clusters(orc):
    n, cp, rv, w, expected = karate()
    for seed in range(5):
        lab, q = orc.louvain(n, cp, rv, w, seed=seed)
        g = orc.group_singletons(n, cp, rv, w, lab)
        assert len(set(g.tolist())) == expected
        assert q == pytest.approx(orc.modularity(n, cp, rv, w, lab), abs=1e-12)


def test_karate_published_optimum(orc):
    """At resolution 1 the karate graph's maximum modularity is 0.4198
    (Brandes et al. 2008, IEEE TKDE 20:172, four communities); the oracle's
    modularity of that partition and its Louvain both reach it."""
    n, cp, rv, w, _ = karate()
    opt = np.zeros(n, np.int64)
    for c, mem in enumerate([[0, 1, 2, 3, 7, 11, 12, 13, 17, 19, 21], [4, 5, 6, 10, 16],
                             [8, 9, 14, 15, 18, 20, 22, 26, 29, 30, 32, 33],
                             [23, 24, 25, 27, 28, 31]]):
        opt[mem] = c
    q = orc.modularity(n, cp, rv, w, opt, 1.0)
    assert round(q, 4) == 0.4198
    A = np.zeros((n, n))
    for j in range(n):
        A[rv[cp[j]:cp[j + 1]], j] = 1.0
    assert q == pytest.approx(dense_modularity(A, opt, 1.0), abs=1e-14)
    _, ql = orc.louvain(n, cp, rv, w, resolution=1.0, seed=1)
    assert ql == pytest.approx(q, abs=1e-12)


@pytest.mark.parametrize("res", [0.5, 0.8, 1.3])
def test_modularity_dense_formula(orc, res):
    A = random_graph(60, 0.15, seed=int(res * 10))
    M = csc(A)
    rng = np.random.default_rng(3)
    lab = rng.integers(0, 7, 60)
    lab = np.unique(lab, return_inverse=True)[1].astype(np.int64)
    q = orc.modularity(60, M.indptr, M.indices, M.data, lab, res)
    assert q == pytest.approx(dense_modularity(A, lab, res), abs=1e-12)


def set_partitions(n):
    """All partitions of range(n) as restricted growth strings."""
    def rec(i, cur, m):
        if i == n:
            yield np.array(cur)
            return
        for c in range(m + 1):
            cur.append(c)
            yield from rec(i + 1, cur, max(m, c + 1))
            cur.pop()
    yield from rec(0, [], 0)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_louvain_vs_brute_force(orc, seed):
    """Eight nodes: Louvain never exceeds the maximum over all 4140
    partitions, and on two weighted cliques joined by one edge it reaches
    it."""
    A = random_graph(8, 0.5, seed=seed)
    M = csc(A)
    best = max(dense_modularity(A, p, 0.8) for p in set_partitions(8))
    lab, q = orc.louvain(8, M.indptr, M.indices, M.data, seed=seed)
    assert q <= best + 1e-12
    assert q >= best - 0.1
    B = np.zeros((8, 8))
    B[:4, :4] = 1.0
    B[4:, 4:] = 1.0
    B[3, 4] = B[4, 3] = 0.2
    np.fill_diagonal(B, 0.0)
    M = csc(B)
    best = max(dense_modularity(B, p, 0.8) for p in set_partitions(8))
    lab, q = orc.louvain(8, M.indptr, M.indices, M.data, seed=seed)
    assert q == pytest.approx(best, abs=1e-12)
    assert lab.tolist() == [0, 0, 0, 0, 1, 1, 1, 1]


def test_group_singletons_rules(orc):
    """Node 6 has entries into clusters 0 and 1 (means 0.5 and 0.9): no
    cluster without an entry, so the larger mean wins.  Node 7 has entries
    into cluster 0 only: cluster 1 has none (mean NaN) and findmax returns
    it although cluster 0 is its only neighbour.  Labels are then renumbered
    by first occurrence."""
    n = 8
    A = np.zeros((n, n))
    for i, j in [(0, 1), (1, 2), (0, 2), (3, 4), (4, 5), (3, 5)]:
        A[i, j] = A[j, i] = 1.0
    A[6, 0] = A[0, 6] = 0.5
    A[6, 3] = A[3, 6] = 0.9
    A[7, 1] = A[1, 7] = 0.4
    np.fill_diagonal(A, 1.0)
    M = csc(A)
    lab = np.array([0, 0, 0, 1, 1, 1, 2, 3])
    g = orc.group_singletons(n, M.indptr, M.indices, M.data, lab)
    assert g.tolist() == [0, 0, 0, 1, 1, 1, 1, 1]
    # every cluster a singleton: the reference's findmax over nothing fails
    assert orc.group_singletons(3, np.array([0, 1, 2, 3]), np.array([0, 1, 2]), np.ones(3),
                                np.array([0, 1, 2])) is None
    # singles move in ascending order: node 7 sees node 6 already in cluster
    # 1, so both clusters have entries and the larger mean (0.4 > 0.3) wins
    lab = np.array([0, 0, 0, 1, 1, 1, 2, 3])
    A2 = A.copy()
    A2[7, 6] = A2[6, 7] = 0.3
    M2 = csc(A2)
    g2 = orc.group_singletons(n, M2.indptr, M2.indices, M2.data, lab)
    assert g2.tolist() == [0, 0, 0, 1, 1, 1, 1, 0]
