"""GPU parity: synthetic counts, normalize_cells, feature gather and
scale_features against the oracle (bit-exact; log1p within 1 ulp)."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def ulp_diff(a, b):
    a = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


@pytest.mark.parametrize("C,G,dens,boost", [(3000, 500, 0.05, 0.0), (4100, 300, 0.04, 0.05),
                                            (1, 7, 0.5, 0.0)])
def test_synth_bit_exact(sev, orc, C, G, dens, boost):
    t = sev.nb_tables(G, seed=11, density=dens, boost_frac=boost)
    p = sev.Pipeline()
    p.synth_counts(C, t)
    X = p.get_counts()
    cp, rv, cnt = orc.synth_csc(C, t)
    assert np.array_equal(X.indptr, cp) and np.array_equal(X.indices, rv)
    assert np.array_equal(X.data, cnt)


@pytest.mark.parametrize("log", [True, False])
def test_pipeline_normalize_vs_oracle(sev, orc, log):
    t = sev.nb_tables(2000, seed=2, density=0.05)
    C = 5000
    p = sev.Pipeline()
    p.synth_counts(C, t)
    cp, rv, cnt = orc.synth_csc(C, t)
    p.normalize("lognormalize" if log else "relativecounts", 1e4)
    got = p.get_normalized()
    ref = orc.normalize(C, cp, rv, cnt, 1e4, log, np.float32)
    d = ulp_diff(got, ref)
    if log:
        assert d.max() <= 1 and (d > 0).mean() < 1e-5
    else:
        assert d.max() == 0
    assert np.all(np.abs(got - ref) <= 1e-5 * np.abs(ref))


@pytest.mark.parametrize("dt,tag", [(np.float32, "f32"), (np.float64, "f64")])
@pytest.mark.parametrize("sf", [1.0, 10.0, 1e4])
@pytest.mark.parametrize("log", [False, True])
def test_normalize_cells_golden(sev, golden, dt, tag, sf, log):
    X = sp.csc_matrix(golden["counts"])
    Y = sev.normalize_cells(X, method="lognormalize" if log else "relativecounts",
                            scale_factor=sf, dtype=dt)
    exp = golden[f"norm_{tag}_sf{sf:g}_{'log' if log else 'rel'}"]
    assert Y.dtype == dt
    if dt == np.float32 and log:
        assert ulp_diff(Y.data, exp).max() <= 1
    else:
        assert np.array_equal(Y.data, exp)


def test_normalize_cells_errors(sev):
    X = sp.csc_matrix(np.array([[1, 0], [0, 2]]))
    with pytest.raises(ValueError):
        sev.normalize_cells(X, method="bogus")


@pytest.mark.parametrize("bad", ["unsorted", "duplicate"])
def test_load_rejects_unsorted_rows(sev, bad):
    """The C-ABI refuses a column whose rows are not strictly increasing
    (a SparseMatrixCSC invariant the row-range searches rely on)."""
    import ctypes as C
    _lib = sev._lib
    colptr = np.array([1, 3, 5], np.int64)  # 1-based, two columns of 2 entries
    rowval = np.array([1, 3, 2, 4] if bad == "unsorted" else [1, 3, 2, 2], np.int64)
    if bad == "unsorted":
        rowval[:2] = [3, 1]
    counts = np.array([1, 2, 3, 4], np.int64)
    out = np.empty(4, np.float64)
    L = _lib.load()
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = L.sev_normalize_cells(4, 2, ptr(colptr), ptr(rowval), ptr(counts), _lib.SEV_INT64,
                               _lib.SEV_LOGNORMALIZE, 1.0, _lib.SEV_FLOAT64, ptr(out))
    assert rc != 0
    assert b"strictly increasing" in L.sev_last_error()


def test_row_totals_many_ranges(sev):
    """Cell totals over several 16384-row ranges and 2048-column chunks
    (ragged at both ends) against scipy's integer row sums."""
    rng = np.random.default_rng(9)
    m, n = 40000, 4500
    X = sp.random(m, n, density=0.004, format="csc", random_state=rng,
                  data_rvs=lambda k: rng.integers(1, 50, k)).astype(np.int64)
    # plus a dense row
    X = sp.csc_matrix(X + sp.csc_matrix((np.full(n, 7), (np.full(n, 5), np.arange(n))),
                                        shape=(m, n)))
    Y = sev.normalize_cells(X, method="relativecounts", scale_factor=1.0, dtype=np.float64)
    tot = np.asarray(X.sum(axis=1)).ravel().astype(np.float64)
    rows = X.indices
    expect = X.data.astype(np.float64) / tot[rows]
    assert np.array_equal(Y.data, expect)


def test_normalize_f64_vs_oracle(sev, orc):
    t = sev.nb_tables(800, seed=5, density=0.08)
    cp, rv, cnt = orc.synth_csc(1500, t)
    X = sp.csc_matrix((cnt, rv, cp), shape=(1500, 800))
    Y = sev.normalize_cells(X, scale_factor=1e4, dtype=np.float64)
    ref = orc.normalize(1500, cp, rv, cnt, 1e4, True, np.float64)
    assert np.allclose(Y.data, ref, rtol=2e-16, atol=0)


@pytest.mark.parametrize("scale_max", [10.0, np.inf, 0.5])
def test_pipeline_scale_bit_exact(sev, orc, scale_max):
    t = sev.nb_tables(3000, seed=8, density=0.05)
    C = 5000
    p = sev.Pipeline()
    p.synth_counts(C, t)
    cp, rv, cnt = orc.synth_csc(C, t)
    p.normalize("lognormalize", 1e4)
    norm = p.get_normalized()
    feats = sev.top_mean_features(t, 500)[::-1].copy()  # non-monotone order
    p.select_scale(feats, scale_max)
    B, mu = p.get_scaled()
    fcp, frow, fval = orc.gather(feats, cp, rv, norm)
    assert np.array_equal(B.indptr, fcp) and np.array_equal(B.indices, frow)
    bref, muref, _ = orc.scale(C, fcp, fval, np.float32(scale_max))
    assert np.array_equal(B.data, bref)
    assert np.array_equal(mu, muref)


def test_scale_features_golden(sev, golden):
    X = sp.csc_matrix(golden["counts"].astype(np.float64))
    S = sev.scale_features(X)
    assert np.allclose(S.mu, golden["scale_mu_over_sd"], rtol=1e-13)
    assert np.allclose(S.to_dense(), golden["scale_dense"], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_scale_features_vs_oracle(sev, orc, dt):
    rng = np.random.default_rng(3)
    m, n = 20000, 40
    D = (rng.random((m, n)) < 0.15) * rng.gamma(2.0, 1.0, (m, n))
    X = sp.csc_matrix(D.astype(dt))
    S = sev.scale_features(X, scale_max=4.0)
    out, mu, _ = orc.scale(m, X.indptr, X.data, dt(4.0))
    assert np.array_equal(S.A.data, out) and np.array_equal(S.mu, mu)


def test_scale_stats_extreme_values(sev, orc):
    """Welford statistics in FP32 on values that stress the quotient: deltas
    below 2^-90 and subnormal ones (the exact-division branch), large
    magnitudes, mixed signs, repeated values and columns of 300,000 stored
    entries (full chunks, counts past 2^16), bit for bit against the
    oracle's FP32 division.  Every column's variance stays a normal float."""
    rng = np.random.default_rng(11)
    m = 300_000
    h = m // 2
    g = rng.gamma(2.0, 1.0, m).astype(np.float32)
    c0 = rng.standard_normal(m).astype(np.float32)
    c1 = g.copy()
    c1[:h] = (np.float32(1e-30) * (1 + rng.integers(0, 4, h))).astype(np.float32)
    c2 = np.where(rng.random(m) < 0.05, g, np.float32(0))
    c2[:1000] = (np.float32(1e-40) * rng.integers(1, 6, 1000) *
                 np.where(rng.random(1000) < 0.5, -1, 1)).astype(np.float32)
    c3 = (rng.standard_normal(m) * 1e15).astype(np.float32)
    c4 = np.where(rng.random(m) < 0.02, np.float32(2.5), np.float32(0))
    c5 = np.where(rng.random(m) < 0.5, np.float32(7.0), np.float32(-7.0))
    X = sp.csc_matrix(np.stack([c0, c1, c2, c3, c4, c5], axis=1))
    assert X.dtype == np.float32 and X.indptr[-1] > 4 * m
    S = sev.scale_features(X, scale_max=np.inf)
    out, mu, _ = orc.scale(m, X.indptr, X.data, np.float32(np.inf))
    assert np.all(np.isfinite(mu))
    assert np.array_equal(S.mu, mu)
    assert np.array_equal(S.A.data, out)


def test_large_load_roundtrip_and_checks(sev):
    """Counts above the pinned-staging threshold (rows 160 MB as Int64,
    values 80 MB) come back unchanged; a row out of range or a column out
    of order anywhere in them is refused."""
    rng = np.random.default_rng(17)
    m, n, per = 400_000, 2000, 10_000  # 20M stored counts
    rows = np.sort(rng.choice(m, size=(n, per), replace=True), axis=1)
    # strictly increasing rows per column: drop repeats column by column
    cols = [np.unique(r) for r in rows]
    colptr = np.concatenate([[0], np.cumsum([len(c) for c in cols])]).astype(np.int64) + 1
    rowval = (np.concatenate(cols).astype(np.int64) + 1)
    vals = rng.integers(1, 50, len(rowval)).astype(np.int32)
    p = sev.Pipeline()
    p.load_counts_julia(m, n, colptr, rowval, vals)
    mm, nn, cp2, rv2, v2 = p.get_counts_julia()
    assert (mm, nn) == (m, n)
    assert np.array_equal(cp2, colptr) and np.array_equal(rv2, rowval)
    assert np.array_equal(v2, vals)
    bad = rowval.copy()
    bad[len(bad) - 3] = 6 + (1 << 32)  # row 5 + 2^32: would wrap to a valid int32 row
    with pytest.raises(sev.SeveroError, match="out of range"):
        p.load_counts_julia(m, n, colptr, bad, vals)
    uns = rowval.copy()
    j = colptr[1500] - 1
    uns[j], uns[j + 1] = uns[j + 1], uns[j]
    with pytest.raises(sev.SeveroError, match="strictly increasing"):
        p.load_counts_julia(m, n, colptr, uns, vals)
    p.close()
