"""GPU parity of find_variable_features(:vst) against the oracle: the
per-gene variances after the loess fit (FP64, within 1e-9 relative: the
device sums run in a different order) and the selected genes."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _counts(sev, orc, cells, genes, density, seed):
    t = sev.nb_tables(genes, seed=seed, density=density)
    cp, rv, cnt = orc.synth_csc(cells, t)
    return t, sp.csc_matrix((cnt, rv, cp), shape=(cells, genes))


@pytest.mark.parametrize("cells,genes,density,nfeat,span",
                         [(5000, 3000, 0.05, 300, 0.5), (800, 1200, 0.1, 1200, 0.3),
                          (3000, 500, 0.02, 50, 0.75)])
def test_vst_matches_oracle(sev, orc, cells, genes, density, nfeat, span):
    _, X = _counts(sev, orc, cells, genes, density, seed=genes)
    sel, metric = sev.find_variable_features(X, nfeat, loess_span=span, return_metric=True)
    rsel, rmetric, mu, sd, esd = orc.vst(cells, X.indptr, X.data, nfeatures=nfeat,
                                         loess_span=span)
    assert np.allclose(metric, rmetric, rtol=1e-9, atol=1e-300)
    assert np.array_equal(metric == 0, rmetric == 0)
    # the selection is the oracle's up to reordering among near ties (genes
    # whose counts differ only in row order have variances 1 ulp apart in
    # either implementation, and the two sum in different orders)
    assert np.allclose(rmetric[sel], rmetric[rsel], rtol=1e-9, atol=0)
    order = np.lexsort((np.arange(genes), -rmetric))
    if nfeat < genes:
        gap = rmetric[order[nfeat - 1]] - rmetric[order[nfeat]]
        assert gap > 1e-8 * abs(rmetric[order[nfeat - 1]])
        assert set(sel.tolist()) == set(rsel.tolist())
    exact = np.array_equal(sel, rsel)
    print(f"selection identical in order: {exact}")


def test_vst_pipeline_equals_host_entry(sev, orc):
    t, X = _counts(sev, orc, 4000, 2000, 0.05, seed=3)
    p = sev.Pipeline()
    p.synth_counts(4000, t)
    sel_p, m_p = p.find_variable_features(200, return_metric=True)
    sel_h, m_h = sev.find_variable_features(p.get_counts(), 200, return_metric=True)
    assert np.array_equal(sel_p, sel_h) and np.array_equal(m_p, m_h)


def test_vst_constant_genes_rank_last(sev):
    rng = np.random.default_rng(9)
    X = rng.poisson(0.5, (600, 40)).astype(np.int32)
    X[:, :5] = 0
    X[:, 5:8] = 2
    sel, metric = sev.find_variable_features(sp.csc_matrix(X), 40, return_metric=True)
    assert np.all(metric[:8] == 0)
    assert set(sel[-8:].tolist()) == set(range(8))


def _oracle_method(orc, X, method, nbins=20, alpha=0.1):
    m = X.shape[0]
    mu, var = orc.rel_mean_var(m, X.indptr, X.indices, X.data)
    if method == "dispersion":
        return orc.hvg_dispersion(mu, var)
    if method == "meanvarplot":
        return orc.hvg_meanvarplot(mu, var, nbins)
    return orc.hvg_saunders(m, X.indptr, X.indices, X.data, mu, var, alpha=alpha)


@pytest.mark.parametrize("method", ["dispersion", "meanvarplot", "saunders"])
@pytest.mark.parametrize("cells,genes,density,nfeat",
                         [(5000, 3000, 0.05, 300), (800, 1200, 0.1, 1200)])
def test_methods_match_oracle(sev, orc, method, cells, genes, density, nfeat):
    """:dispersion, :meanvarplot (20 width bins) and :saunders (alpha 0.1):
    the per-gene metric bit for bit (the device Welford of x / s runs in the
    oracle's order) and the selected genes in order."""
    _, X = _counts(sev, orc, cells, genes, density, seed=genes + 7)
    X = sp.csc_matrix(X)
    X.sort_indices()
    sel, metric = sev.find_variable_features(X, nfeat, method=method, return_metric=True)
    rmetric = _oracle_method(orc, X, method)
    assert np.array_equal(metric, rmetric)
    assert np.array_equal(sel, orc.top_features(rmetric, nfeat))


def test_meanvarplot_bins_and_errors(sev, orc):
    """Other bin counts; frequency binning and unknown methods are refused."""
    _, X = _counts(sev, orc, 2000, 800, 0.08, seed=31)
    X = sp.csc_matrix(X)
    for nb in (5, 37):
        _, metric = sev.find_variable_features(X, 100, method="meanvarplot", num_bins=nb,
                                               return_metric=True)
        assert np.array_equal(metric, _oracle_method(orc, X, "meanvarplot", nbins=nb))
    with pytest.raises(ValueError):
        sev.find_variable_features(X, 100, method="meanvarplot", binning_method="frequency")
    with pytest.raises(ValueError):
        sev.find_variable_features(X, 100, method="bogus")


@pytest.mark.parametrize("method", ["dispersion", "meanvarplot", "saunders"])
def test_methods_pipeline_equals_host_entry(sev, orc, method):
    """The pipeline's selection on its resident counts is the one-shot
    entry's on the same counts, bit for bit."""
    t, _ = _counts(sev, orc, 3000, 1500, 0.05, seed=17)
    p = sev.Pipeline()
    p.synth_counts(3000, t)
    sel_p, m_p = p.find_variable_features(150, method=method, return_metric=True)
    sel_h, m_h = sev.find_variable_features(p.get_counts(), 150, method=method,
                                            return_metric=True)
    assert np.array_equal(sel_p, sel_h) and np.array_equal(m_p, m_h)
    with pytest.raises(ValueError):
        p.find_variable_features(10, method="bogus")
