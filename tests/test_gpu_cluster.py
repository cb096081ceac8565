"""GPU Louvain clustering (cluster(SNN), src/clustering.jl:59-104,
src/modularity.jl:165-380) against the oracle's sequential restatement.

The reference's local moving is a randomised sequential search and the
device moves nodes in parallel sweeps, so labels are compared through what
fixes them: the karate fixture of test/test_cluster.jl (3 clusters),
planted partitions recovered exactly, the reported modularity against the
oracle's modularity of the same labels (1e-9: weights are summed in fixed
point with 2^-40 resolution), the modularity reached against the oracle's
best start (within 0.01), and group_singletons! exactly as the oracle on
the same ungrouped labels."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
from sklearn.metrics import adjusted_rand_score

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def karate():
    d = json.load(open(os.path.join(HERE, "golden", "karate.json")))
    n = d["n"]
    M = sp.csc_matrix((np.ones(len(d["indices"])), d["indices"], d["indptr"]), shape=(n, n))
    return M, d["expected_nclusters"]


def planted(sizes, p_in, p_out, seed, weighted=True, self_links=True):
    rng = np.random.default_rng(seed)
    n = sum(sizes)
    block = np.repeat(np.arange(len(sizes)), sizes)
    P = np.where(block[:, None] == block[None, :], p_in, p_out)
    U = np.triu(rng.random((n, n)) < P, 1).astype(np.float64)
    if weighted:
        U *= rng.uniform(0.1, 1.0, (n, n))
    A = U + U.T
    if self_links:
        np.fill_diagonal(A, 1.0)
    M = sp.csc_matrix(A)
    M.sort_indices()
    return M, block


def oracle_cluster(orc, M, **kw):
    M = sp.csc_matrix(M)
    M.sort_indices()
    return orc.louvain(M.shape[0], M.indptr, M.indices, M.data, **kw)


def first_occurrence(lab):
    seen = {}
    for x in lab:
        seen.setdefault(x, len(seen) + 1)
    return np.array([seen[x] for x in lab])


def test_karate_fixture(sev, orc):
    M, expected = karate()
    lab, q = sev.cluster(M, return_modularity=True)
    assert len(set(lab.tolist())) == expected
    assert np.array_equal(lab, first_occurrence(lab))
    _, qo = oracle_cluster(orc, M)
    print(f"karate modularity device {q:.6f} oracle {qo:.6f}")
    assert q >= qo - 0.01


@pytest.mark.parametrize("sizes,p_in,p_out,seed",
                         [([400] * 8, 0.08, 0.002, 1), ([400] * 4, 0.4, 0.01, 4),
                          ([500] * 4, 0.6, 0.02, 3)])
def test_planted_partition(sev, orc, sizes, p_in, p_out, seed):
    """Well separated blocks are recovered exactly by both.  Node degrees:
    about 35 (256-slot LDS tables), about 170 (512-slot tables of the
    listed launch) and about 330 (a block each, global tables)."""
    M, block = planted(sizes, p_in, p_out, seed)
    lab, q = sev.cluster(M, return_modularity=True, group_singletons=False)
    assert adjusted_rand_score(block, lab) == 1.0
    lo, qo = oracle_cluster(orc, M, nstarts=3)
    assert adjusted_rand_score(block, lo) == 1.0
    assert q == pytest.approx(qo, abs=1e-9)
    assert q == pytest.approx(orc.modularity(M.shape[0], M.indptr, M.indices, M.data, lab - 1),
                              abs=1e-9)


def test_resolution_limit_merge(sev, orc):
    """Unequal blocks: at resolution 0.8 modularity prefers merging the small
    ones (the planted partition scores lower), and the device reaches the
    oracle's modularity within 0.005 with nearly the same partition."""
    M, block = planted([50, 300, 120, 700, 90], 0.1, 0.003, seed=2)
    lab, q = sev.cluster(M, return_modularity=True, group_singletons=False)
    lo, qo = oracle_cluster(orc, M)
    qp = orc.modularity(M.shape[0], M.indptr, M.indices, M.data, block)
    ari = adjusted_rand_score(lo, lab)
    print(f"modularity device {q:.6f} oracle {qo:.6f} planted {qp:.6f}; ARI {ari:.4f}")
    assert qo > qp
    assert q >= qo - 0.005
    assert ari > 0.95


def test_hub_node(sev, orc):
    """A hub joined to every node of ten cliques (degree 500) takes the
    block-per-node path; the cliques are still found."""
    n = 501
    A = np.zeros((n, n))
    for c in range(10):
        s = slice(1 + 50 * c, 1 + 50 * (c + 1))
        A[s, s] = 1.0
    A[0, 1:] = A[1:, 0] = 0.05
    np.fill_diagonal(A, 1.0)
    M = sp.csc_matrix(A)
    lab, q = sev.cluster(M, return_modularity=True, group_singletons=False)
    block = np.concatenate([[0], np.repeat(np.arange(10), 50)])
    assert adjusted_rand_score(block[1:], lab[1:]) == 1.0
    assert q == pytest.approx(orc.modularity(n, M.indptr, M.indices, M.data, lab - 1), abs=1e-9)
    _, qo = oracle_cluster(orc, M, nstarts=3)
    assert q >= qo - 1e-9


@pytest.mark.parametrize("n,centers,seed", [(6000, 12, 5), (20000, 30, 6)])
def test_snn_graph_vs_oracle(sev, orc, n, centers, seed):
    """The Jaccard SNN of a kNN graph (k = 20, prune 1/15) of a Gaussian
    mixture: the modularity reached is within 0.01 of the oracle's best of
    10 sequential starts, and the partitions agree (ARI > 0.9)."""
    rng = np.random.default_rng(seed)
    C = rng.normal(0, 6.0, (centers, 16))
    X = (C[rng.integers(0, centers, n)] + rng.normal(0, 1.0, (n, 16))).astype(np.float32)
    snn = sev.shared_nearest_neighbours(X, 20)
    lab, q = sev.cluster(snn, return_modularity=True, group_singletons=False)
    lo, qo = oracle_cluster(orc, snn)
    ari = adjusted_rand_score(lo, lab)
    print(f"n={n}: modularity device {q:.6f} oracle {qo:.6f}; clusters {lab.max()} / "
          f"{lo.max() + 1}; ARI {ari:.4f}")
    assert q == pytest.approx(orc.modularity(n, snn.indptr, snn.indices,
                                             snn.data.astype(np.float64), lab - 1), abs=1e-9)
    assert q >= qo - 0.01
    assert ari > 0.9


def test_deterministic_and_seeded(sev):
    M, _ = planted([300] * 6, 0.05, 0.01, seed=9)
    a = sev.cluster(M, seed=4)
    b = sev.cluster(M, seed=4)
    assert np.array_equal(a, b)
    assert np.array_equal(a, first_occurrence(a))


def test_singletons_grouped_as_oracle(sev, orc):
    """Isolated nodes stay alone after Louvain; group_singletons! then puts
    each in the first cluster it has no entry into (the reference's NaN
    mean), exactly as the oracle does from the same ungrouped labels."""
    M, _ = planted([40, 60, 30], 0.5, 0.01, seed=11)
    A = M.toarray()
    n0 = A.shape[0]
    B = np.zeros((n0 + 3, n0 + 3))
    B[:n0, :n0] = A
    B[n0:, n0:] = np.eye(3)
    order = np.random.default_rng(2).permutation(n0 + 3)
    B = B[np.ix_(order, order)]
    Ms = sp.csc_matrix(B)
    Ms.sort_indices()
    ungrouped = sev.cluster(Ms, group_singletons=False)
    counts = np.bincount(ungrouped)
    assert (counts[1:] == 1).sum() == 3
    grouped = sev.cluster(Ms)
    ref = orc.group_singletons(Ms.shape[0], Ms.indptr, Ms.indices, Ms.data, ungrouped - 1)
    assert np.array_equal(grouped - 1, ref)


def test_no_edges_and_errors(sev):
    n = 5
    M = sp.identity(n, format="csc")
    lab, q = sev.cluster(M, group_singletons=False, return_modularity=True)
    assert lab.tolist() == [1, 2, 3, 4, 5] and np.isnan(q)
    with pytest.raises(sev.SeveroError):
        sev.cluster(M)
    with pytest.raises(ValueError):
        sev.cluster(sp.csc_matrix(np.ones((3, 4))))
    with pytest.raises(ValueError):
        sev.cluster(M, algorithm="leiden")
    bad = sp.csc_matrix(np.array([[0.0, -1.0], [-1.0, 0.0]]))
    with pytest.raises(sev.SeveroError):
        sev.cluster(bad)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_pipeline_cluster_equals_one_shot(sev, orc, dtype):
    """Clustering the device-resident SNN gives the one-shot entry's labels
    on the same graph, bit for bit."""
    rng = np.random.default_rng(31)
    C = rng.normal(0, 5.0, (9, 30))
    Z = np.asfortranarray((C[rng.integers(0, 9, 4000)] + rng.normal(0, 1.0, (4000, 30)))
                          .astype(np.float32))
    p = sev.Pipeline()
    p.set_embedding(Z)
    p.neighbours(20, include_self=True, ndims=30)
    p.snn(1 / 15, dtype)
    lab_p, q_p = p.cluster(4000, return_modularity=True)
    S = p.get_snn(4000)
    p.close()
    lab_h, q_h = sev.cluster(S, return_modularity=True)
    assert np.array_equal(lab_p, lab_h) and q_p == q_h
    assert lab_p.max() == 9


def test_dependent_singletons_in_order(sev, orc):
    """Two singles joined by a stored zero-weight entry (weight 0 gives them
    no reason to move): the first goes to the first cluster, and the second,
    which then has an entry into that cluster, to the next one, as the
    reference's in-order loop places them."""
    M, _ = planted([40, 60, 30], 0.5, 0.01, seed=12)
    A = M.toarray()
    n0 = A.shape[0]
    n = n0 + 4
    B = np.zeros((n, n))
    B[:n0, :n0] = A
    B[np.arange(n0, n), np.arange(n0, n)] = 1.0
    cols, rows, vals = [], [], []
    for j in range(n):
        r = list(np.nonzero(B[:, j])[0])
        if j in (n0 + 1, n0 + 2):  # the stored zero between nodes n0+1 and n0+2
            r.append(n0 + 2 if j == n0 + 1 else n0 + 1)
        r.sort()
        rows += r
        vals += [B[i, j] for i in r]
        cols.append(len(r))
    indptr = np.concatenate([[0], np.cumsum(cols)])
    Ms = sp.csc_matrix((np.array(vals), np.array(rows), indptr), shape=(n, n))
    assert Ms.nnz == np.count_nonzero(B) + 2
    ungrouped = sev.cluster(Ms, group_singletons=False)
    grouped = sev.cluster(Ms)
    ref = orc.group_singletons(n, Ms.indptr, Ms.indices, Ms.data, ungrouped - 1)
    assert np.array_equal(grouped - 1, ref)
    assert grouped[n0 + 1] != grouped[n0 + 2]


def test_full_size_properties(sev, orc):
    """1,000,000 cells (50 Gaussian centres in 16 dimensions, Jaccard SNN of
    k = 20, about 33M entries), one start: the reported modularity equals the
    oracle's modularity of the returned labels (1e-9), the labels are in
    order of first occurrence, and the centres are recovered (ARI > 0.99)."""
    n = 1_000_000
    rng = np.random.default_rng(1)
    C = rng.normal(0, 6.0, (50, 16))
    centre = rng.integers(0, 50, n)
    X = (C[centre] + rng.normal(0, 1.0, (n, 16))).astype(np.float32)
    snn = sev.shared_nearest_neighbours(X, 20)
    del X
    lab, q = sev.cluster(snn, nrandomstarts=1, return_modularity=True, group_singletons=False)
    assert lab[0] == 1 and np.all(np.diff(np.maximum.accumulate(lab)) <= 1)
    qo = orc.modularity(n, snn.indptr, snn.indices, snn.data.astype(np.float64), lab - 1)
    print(f"1M cells: {lab.max()} clusters, modularity {q:.6f} (oracle on the labels {qo:.6f})")
    assert q == pytest.approx(qo, abs=1e-9)
    assert adjusted_rand_score(centre, lab) > 0.99
