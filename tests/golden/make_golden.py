"""Writes tests/golden/golden.npz.

Inputs are the fixtures the reference's own tests hold:
  * the 10 x 5 count matrix of test/test_input.jl:47-58 (= test/test_scaling.jl:22-33),
  * the CenteredMatrix known-answer vectors of test/test_scaling.jl:72-111.
Expected outputs are computed here from closed forms with numpy (IEEE float32 /
float64 elementwise arithmetic, no FMA), not from any implementation under test:
  * normalize: v = (sf * x) / s_cell in the working precision, left to right
    (src/normalize.jl:27), log1p for :lognormalize, with FP32 log1p defined as
    the double log1p rounded to float;
  * scale: (X - mean) / std with the sample std (test/test_scaling.jl:41-44);
  * centered products: the dense Q = A - 1 mu' (test/test_scaling.jl:74-91).
Run:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

COUNTS = np.array([
    [0, 0, 0, 3, 0],
    [0, 1, 0, 0, 6],
    [5, 0, 0, 0, 0],
    [3, 0, 0, 0, 0],
    [0, 0, 6, 2, 0],
    [0, 0, 0, 0, 0],
    [0, 0, 0, 0, 2],
    [0, 3, 0, 0, 0],
    [0, 0, 0, 0, 3],
    [2, 0, 0, 0, 0],
], dtype=np.int64)


def csc_parts(D):
    """0-based CSC of a dense matrix, rows ascending within each column."""
    colptr = [0]
    rows, vals = [], []
    for j in range(D.shape[1]):
        nz = np.nonzero(D[:, j])[0]
        rows.extend(nz.tolist())
        vals.extend(D[nz, j].tolist())
        colptr.append(len(rows))
    return np.array(colptr, np.int64), np.array(rows, np.int32), np.array(vals)


def normalize_expected(D, sf, log, dt):
    colptr, rows, vals = csc_parts(D)
    tot = D.sum(axis=1)
    out = np.empty(len(vals), dt)
    for p in range(len(vals)):
        v = (dt(sf) * dt(vals[p])) / dt(tot[rows[p]])
        if log:
            v = dt(np.log1p(np.float64(v)))
        out[p] = v
    return out


def main():
    g = {"counts": COUNTS}
    colptr, rows, vals = csc_parts(COUNTS)
    g["colptr"], g["rowval"], g["nzval"] = colptr, rows, vals.astype(np.int32)
    for dt, tag in ((np.float32, "f32"), (np.float64, "f64")):
        for sf in (1.0, 10.0, 1e4):
            for log in (False, True):
                key = f"norm_{tag}_sf{sf:g}_{'log' if log else 'rel'}"
                g[key] = normalize_expected(COUNTS, sf, log, dt)
    X = COUNTS.astype(np.float64)
    mu = X.mean(axis=0)
    sd = X.std(axis=0, ddof=1)
    g["scale_mu_over_sd"] = mu / sd
    g["scale_dense"] = (X - mu) / sd
    A = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1]], np.float64)
    cmu = np.array([0.1, 0.2, 0.3])
    r = np.array([0.991, 0.228, 0.291])
    y0 = np.array([0.820, 0.430, 0.264, 0.789])
    Q = A - cmu[None, :]
    y1 = 2 * Q @ r + y0
    g["cm_A"], g["cm_mu"], g["cm_r"], g["cm_y0"] = A, cmu, r, y0
    g["cm_Qr"] = Q @ r
    g["cm_y1"] = y1
    g["cm_Qty1"] = Q.T @ y1
    g["cm_r1"] = 2 * Q.T @ y1 + r
    np.savez(os.path.join(HERE, "golden.npz"), **g)
    print("wrote", os.path.join(HERE, "golden.npz"), len(g), "arrays")


if __name__ == "__main__":
    main()
