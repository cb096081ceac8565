"""Synthetic single-cell count matrices (negative binomial, seeded).

The generator itself runs on the device (``sev_synth_nb_csc`` in the C-ABI).
This module computes its small per-gene host tables:

* per-gene mean ``m_g = exp(N(mu, sigma^2))`` (log-normal, numpy seeded),
* NB size ``r = 1/dispersion``,
* a base per-cell size factor ``s`` calibrated so the expected density of
  nonzeros matches the request (``boost_frac`` of the cells get ``boost``
  times the library size),
* per-(gene, class) tables ``p0 = (r/(r+mu))^r`` (probability of a zero)
  and ``q = mu/(r+mu)``, flattened as index ``2*g + class``.

Each entry (gene g, cell c) is drawn by inverse CDF from one counter-based
uniform, using only IEEE ``+ * /`` on doubles, so every device and host
implementation produces the same matrix bit for bit.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class NBTables:
    n_genes: int
    seed: int
    r: float
    size_factor: float
    boost_frac: float
    boost: float
    gene_mean: np.ndarray  # (G,) float64
    p0: np.ndarray  # (2G,) float64
    q: np.ndarray  # (2G,) float64

    def expected_density(self) -> float:
        d0 = 1.0 - self.p0[0::2]
        d1 = 1.0 - self.p0[1::2]
        return float((1.0 - self.boost_frac) * d0.mean() + self.boost_frac * d1.mean())


def _p0_q(mu: np.ndarray, r: float):
    p0 = np.power(r / (r + mu), r)
    q = mu / (r + mu)
    return p0, q


def nb_tables(n_genes: int, seed: int = 0, density: float = 0.05, log_mean: float = -1.0,
              log_sd: float = 1.5, dispersion: float = 0.3, boost_frac: float = 0.0,
              boost: float = 3.0) -> NBTables:
    """Host tables for the seeded negative-binomial generator.

    ``density`` is the target expected fraction of nonzeros; the base size
    factor is found by bisection on log(s).
    """
    rng = np.random.default_rng(seed)
    m = np.exp(rng.normal(log_mean, log_sd, n_genes))
    r = 1.0 / dispersion

    def dens(s: float) -> float:
        p0a, _ = _p0_q(s * m, r)
        p0b, _ = _p0_q(s * boost * m, r)
        return float((1.0 - boost_frac) * (1.0 - p0a).mean() + boost_frac * (1.0 - p0b).mean())

    lo, hi = -20.0, 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if dens(np.exp(mid)) < density:
            lo = mid
        else:
            hi = mid
    s = float(np.exp(0.5 * (lo + hi)))
    p0 = np.empty(2 * n_genes)
    q = np.empty(2 * n_genes)
    p0[0::2], q[0::2] = _p0_q(s * m, r)
    p0[1::2], q[1::2] = _p0_q(s * boost * m, r)
    return NBTables(n_genes, seed, r, s, boost_frac, boost, m, np.ascontiguousarray(p0),
                    np.ascontiguousarray(q))


def top_mean_features(t: NBTables, nfeatures: int) -> np.ndarray:
    """0-based indices of the ``nfeatures`` genes with the largest generator
    mean, largest first (ties by lower index).  Used as the feature list for
    synthetic workloads in place of a variable-feature ranking."""
    order = np.lexsort((np.arange(t.n_genes), -t.gene_mean))
    return np.ascontiguousarray(order[:nfeatures].astype(np.int64))
