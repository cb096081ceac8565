"""GPU parity at full scale: 4,000,000 points, k = 30 (the large precision
workload), every candidate form checked bit for bit against the oracle's
FP32 brute force on a seeded sample of query rows."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, D, K, NQ = 4_000_000, 50, 30, 200


@pytest.fixture(scope="module")
def large(sev):
    rng = np.random.default_rng(11)
    centers = (rng.standard_normal((400, D)) * 4.0).astype(np.float32)
    Z = centers[rng.integers(0, len(centers), N)]
    Z += rng.standard_normal((N, D), dtype=np.float32)
    Z = np.asfortranarray(Z)
    p = sev.Pipeline()
    p.set_embedding(Z)
    q = np.sort(rng.choice(N, NQ, replace=False))
    return dict(p=p, Z=Z, q=q)


@pytest.mark.parametrize("precision", ["fp8", "bf16"])
def test_knn_4m_sampled_rows_bit_exact(large, orc, precision):
    p, Z, q = large["p"], large["Z"], large["q"]
    nfb = p.neighbours(K, include_self=True, precision=precision)
    idx, dist = p.get_neighbours(N)
    ridx, rdist = orc.knn(Z, K, qidx=q)
    assert np.array_equal(idx[q], ridx)
    assert np.array_equal(dist[q], rdist)
    print(f"{precision}: {nfb} rows fell back to the exhaustive kernel")
    if precision == "fp8":
        left = p.knn_passes()
        print(f"rows left after FP8 / BF16 / FP32 candidates: {left.tolist()}")
    # SNN of the full graph: symmetric, unit diagonal, and the sampled
    # columns' weights from their own neighbour sets
    p.snn(1 / 15, np.float32)
    S = p.get_snn(N)
    assert np.allclose(S.diagonal(), 1.0)
    Sq = S[:, q].toarray()
    assert np.array_equal(Sq, S[q, :].toarray().T)
    sets = {int(i): set(idx[i].tolist()) for i in q}
    for a, i in enumerate(q[:20]):
        for j in q[:20]:
            x = len(sets[int(i)] & sets[int(j)])
            w = np.float32(x) / (np.float32(K) + (np.float32(K) - np.float32(x)))
            expect = w if w > np.float32(1 / 15) else np.float32(0)
            assert Sq[j, a] == expect
