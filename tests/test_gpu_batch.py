"""GPU parity of the batch form: samples of uneven size stacked by rows run
through normalise -> per-sample features and scaling -> per-sample IRLBA ->
per-sample kNN -> SNN, each stage checked against the oracle run on every
sample on its own."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

CELLS = [300, 1200, 650, 2500, 480]
DENS = [0.03, 0.08, 0.05, 0.02, 0.06]
G, F, NPCS, K = 3000, 300, 20, 10


def _bf16(x):
    """float32 -> bfloat16 (round to nearest even) -> float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


@pytest.fixture(scope="module")
def batch(sev, orc):
    S = len(CELLS)
    tabs = [sev.nb_tables(G, seed=100 + s, density=DENS[s]) for s in range(S)]
    feats = np.stack([sev.top_mean_features(t, F) for t in tabs])
    row0 = np.concatenate([[0], np.cumsum(CELLS)])
    ref = []
    for s in range(S):
        cp, rv, cnt = orc.synth_csc(CELLS[s], tabs[s])
        nv = orc.normalize(CELLS[s], cp, rv, cnt, 1e4)
        fcp, frow, fval = orc.gather(feats[s], cp, rv, nv)
        B, mu, sd = orc.scale(CELLS[s], fcp, fval, 10.0)
        ref.append(dict(cp=cp, rv=rv, cnt=cnt, fcp=fcp, frow=frow, B=B, mu=mu))
    p = sev.Pipeline()
    p.synth_batch(CELLS, tabs)
    p.normalize("lognormalize", 1e4)
    p.select_scale_batch(feats, 10.0)
    rng = np.random.default_rng(7)
    init = rng.standard_normal((S, F))
    it, nm = p.pca_batch(NPCS, init=init, tol=1e-5, graph=True)
    pca = p.get_pca_batch()
    return dict(p=p, tabs=tabs, feats=feats, row0=row0, ref=ref, init=init, it=it, pca=pca)


def test_batch_counts_are_stacked_samples(batch):
    X = batch["p"].get_counts()
    Y = sp.vstack([sp.csc_matrix((r["cnt"], r["rv"], r["cp"]), shape=(c, G))
                   for r, c in zip(batch["ref"], CELLS)]).tocsc()
    Y.sort_indices()
    assert np.array_equal(X.indptr, Y.indptr)
    assert np.array_equal(X.indices, Y.indices)
    assert np.array_equal(X.data, Y.data)


def test_batch_scale_per_sample_bitexact(batch):
    Bs, mu = batch["p"].get_scaled()
    assert mu.shape == (len(CELLS), F)
    row0 = batch["row0"]
    for s, r in enumerate(batch["ref"]):
        Bsub = Bs[row0[s]:row0[s + 1]].tocsc()
        assert np.array_equal(Bsub.indptr, r["fcp"]), s
        assert np.array_equal(Bsub.indices, r["frow"]), s
        assert np.array_equal(Bsub.data, r["B"]), s
        assert np.array_equal(mu[s], r["mu"]), s


def test_batch_pca_per_sample(batch, orc):
    """Each sample's triplets against the oracle's FP64 IRLBA on the same
    BF16-rounded values: singular values within 1e-5 (relative), singular
    vectors of well-separated components within 1e-3 rad."""
    s_all, U, V, Z = batch["pca"]
    row0 = batch["row0"]
    assert s_all.shape == (len(CELLS), NPCS) and V.shape == (len(CELLS), F, NPCS)
    for s, r in enumerate(batch["ref"]):
        m = CELLS[s]
        rc, so, Uo, Vo, _, _ = orc.irlba_csc(m, r["fcp"], r["frow"], _bf16(r["B"]), r["mu"],
                                             NPCS, batch["init"][s])
        assert rc == 0
        assert np.allclose(s_all[s], so, rtol=1e-5, atol=0), (s, s_all[s], so)
        gap = np.full(NPCS, np.inf)
        gap[:-1] = np.minimum(gap[:-1], np.diff(-so) / so[:-1])
        gap[1:] = np.minimum(gap[1:], np.diff(-so) / so[1:])
        for i in np.nonzero(gap > 1e-2)[0]:
            cv = abs(V[s][:, i] @ Vo[:, i])
            cu = abs(U[row0[s]:row0[s + 1], i] @ Uo[:, i])
            assert np.arccos(min(cv, 1.0)) < 1e-3, (s, i, cv)
            assert np.arccos(min(cu, 1.0)) < 1e-3, (s, i, cu)
        # the embedding is U diag(s) in Float32
        Zs = (U[row0[s]:row0[s + 1]] * s_all[s]).astype(np.float32)
        assert np.array_equal(Z[row0[s]:row0[s + 1]], Zs)


def test_batch_pca_graph_replay_identical(batch, sev):
    """The round sequence replayed from HIP graphs and launched directly give
    the same bits."""
    p = batch["p"]
    it, _ = p.pca_batch(NPCS, init=batch["init"], tol=1e-5, graph=False)
    s2, U2, V2, Z2 = p.get_pca_batch()
    s1, U1, V1, Z1 = batch["pca"]
    assert np.array_equal(it, batch["it"])
    assert np.array_equal(s1, s2) and np.array_equal(U1, U2) and np.array_equal(V1, V2)
    # a second graph run (cached graphs) again gives the same bits
    p.pca_batch(NPCS, init=batch["init"], tol=1e-5, graph=True)
    s3, U3, _, _ = p.get_pca_batch()
    assert np.array_equal(s1, s3) and np.array_equal(U1, U3)


@pytest.mark.parametrize("precision", ["fp32", "bf16", "exact"])
@pytest.mark.parametrize("include_self", [True, False])
def test_batch_knn_snn_per_sample(batch, sev, orc, precision, include_self):
    p = batch["p"]
    Z = batch["pca"][3]
    p.set_embedding(Z)  # same rows: the batch stays declared
    M = int(sum(CELLS))
    p.neighbours(K, include_self=include_self, precision=precision)
    idx, dist = p.get_neighbours(M)
    row0 = batch["row0"]
    blocks = []
    for s in range(len(CELLS)):
        ri, rd = orc.knn(np.asfortranarray(Z[row0[s]:row0[s + 1]]), K,
                         include_self=include_self)
        assert np.array_equal(idx[row0[s]:row0[s + 1]], ri + row0[s]), (s, precision)
        assert np.array_equal(dist[row0[s]:row0[s + 1]], rd), (s, precision)
        cp, rv, nz = orc.snn(ri, K, 1 / 15, np.float32)
        blocks.append(sp.csc_matrix((nz, rv - 1, cp - 1), shape=(CELLS[s], CELLS[s])))
    p.snn(1 / 15, np.float32)
    Sg = p.get_snn(M)
    Sr = sp.block_diag(blocks, format="csc")
    Sr.sort_indices()
    assert np.array_equal(Sg.indptr, Sr.indptr)
    assert np.array_equal(Sg.indices, Sr.indices)
    assert np.array_equal(Sg.data, Sr.data)


@pytest.mark.parametrize("precision", ["fp32", "bf16", "exact"])
def test_batch_knn_cosine_per_sample(batch, orc, precision):
    """Cosine neighbours of a batch: each sample's lists are the oracle's
    cosine search on that sample alone."""
    p = batch["p"]
    Z = batch["pca"][3]
    p.set_embedding(Z)
    M = int(sum(CELLS))
    p.neighbours(K, include_self=True, precision=precision, metric="cosine")
    idx, dist = p.get_neighbours(M)
    row0 = batch["row0"]
    for s in range(len(CELLS)):
        ri, rd = orc.knn_cosine(np.asfortranarray(Z[row0[s]:row0[s + 1]]), K)
        assert np.array_equal(idx[row0[s]:row0[s + 1]], ri + row0[s]), (s, precision)
        assert np.array_equal(dist[row0[s]:row0[s + 1]], rd), (s, precision)


def test_batch_pca_wider_basis(batch, orc, sev):
    """work = 48 Lanczos vectors per sample: the same singular values as the
    oracle's IRLBA with the reference's npcs + 7, in fewer restarts; a basis
    outside [npcs + 1, 64] is refused."""
    p = batch["p"]
    it, _ = p.pca_batch(NPCS, init=batch["init"], tol=1e-5, graph=True, work=48)
    s_all = p.get_pca_batch()[0]
    assert it.max() < batch["it"].max(), (it, batch["it"])
    for s, r in enumerate(batch["ref"]):
        rc, so, _, _, _, _ = orc.irlba_csc(CELLS[s], r["fcp"], r["frow"], _bf16(r["B"]), r["mu"],
                                           NPCS, batch["init"][s])
        assert rc == 0
        assert np.allclose(s_all[s], so, rtol=1e-5, atol=0), (s, s_all[s], so)
    for bad in (NPCS, 65):
        with pytest.raises(sev.SeveroError, match="batch work"):
            p.pca_batch(NPCS, init=batch["init"], tol=1e-5, work=bad)
