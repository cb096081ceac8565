"""Throughput of Severo's embedding-and-graph path on one MI355X.

A step is one pass of normalize_cells -> Y[:, features] -> scale_features
(clip 10) -> PCA (IRLBA, 50 components, tol 1e-5) -> exact kNN (k = 20) ->
Jaccard SNN (prune 1/15) over a synthetic negative-binomial count matrix that
is generated on the device before timing starts (inputs resident in HBM).

Workloads (cells x genes, density; features = the 2000 genes that the
device's find_variable_features(:vst) selects, computed once before timing;
--features top-mean takes the 2000 genes of largest generator mean instead):
  small : 5,000 x 20,000, 5%
  100k  : 100,000 x 33,538, 4%, 5% of cells with 3x library size
  1m    : 1,000,000 x 27,998, 3%   (default: the quoted metric's workload)
  4m    : 4,000,000 x 2,000, 12%, all 2000 genes as features, k = 30, forms
          bf16, fp8 (FP8 e4m3 kNN candidates) and fp32, IRLBA restart limit 5000

The headline value is the precision form (BF16 operator values for the
Lanczos products, BF16 candidate search for kNN with exact FP32 re-rank);
the plain FP32 form is timed in the same run.  IRLBA keeps --work Lanczos
vectors (default 128; the reference's npcs + 7 = 57 is timed beside it in
the BF16 form, and so is the top-mean feature set).  The batching form (64 samples
of uneven size, 20,000 genes, 30 PCs, k = 15) is timed after it, against the
single-sample pipeline looped over the same samples.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import __graft_entry__ as entry  # noqa: E402

WORKLOADS = {
    "small": dict(cells=5000, genes=20000, density=0.05, boost=0.0),
    "100k": dict(cells=100_000, genes=33538, density=0.04, boost=0.05),
    "1m": dict(cells=1_000_000, genes=27998, density=0.03, boost=0.0),
    # an unstructured NB matrix has a flat spectrum: with the reference's
    # npcs + 7 vectors its 50-component IRLBA needs about 2400 restarts, so
    # the restart limit is raised
    "4m": dict(cells=4_000_000, genes=2000, density=0.12, boost=0.0, k=30,
               forms="bf16,fp32", maxit=5000),
}
NFEAT, NPCS, K, PRUNE, SCALE_MAX, SF = 2000, 50, 20, 1.0 / 15, 10.0, 1e4
MAXIT = 1000  # IRLBA restarts (the reference's default)
DEFAULT_WORK = 112  # Lanczos vectors kept by IRLBA (0: npcs + 7, the reference's choice)
WORK = DEFAULT_WORK
PEAK_HBM_GBS = 8000.0
PEAK_TFLOPS = {"fp32": 157.3, "bf16": 2500.0, "fp8": 5000.0}


def run_step(p, feats, init, form, work=None):
    """One step; the fp8 form is the bf16 form with FP8 kNN candidates."""
    p.normalize("lognormalize", SF)
    p.select_scale(feats, SCALE_MAX)
    it, nm = p.pca(NPCS, init=init, tol=1e-5, maxit=MAXIT,
                   precision="bf16" if form == "fp8" else form,
                   work=WORK if work is None else work)
    nfb = p.neighbours(K, include_self=True, precision=form)
    nnz = p.snn(PRUNE, np.float32)
    p.synchronize()
    return it, nm, nfb, nnz


def time_form(p, feats, init, form, steps, warmup, profile):
    """Warm-up steps, one untimed step with every region timed (the phase
    breakdown), then the timed steps with events around the dominant
    region only (its live launch durations give the roofline figure; the
    other regions add no events to the timed stream)."""
    for _ in range(warmup):
        run_step(p, feats, init, form)
    breakdown, prof = {}, {}
    if profile:
        p.profile(True)
        run_step(p, feats, init, form)
        breakdown = p.profile_read()
        p.profile(True, only=max(breakdown.items(), key=lambda kv: kv[1][1])[0])
    p.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        it, nm, nfb, nnz = run_step(p, feats, init, form)
    p.synchronize()
    dt = (time.perf_counter() - t0) / steps
    if profile:
        prof = p.profile_read()
        p.profile(False)
    info = dict(irlba_iters=it, irlba_nmult=nm, knn_fallback_rows=nfb, snn_nnz=nnz)
    if form == "fp8":
        info["knn_rows_left_after_fp8_bf16_fp32"] = [int(x) for x in p.knn_passes()]
    return dt, info, prof, breakdown


def time_variant(p, feats, init, form, steps, work=None):
    """A secondary timing (no profile): one warm-up step, then `steps`.
    An IRLBA that reaches maxit is reported, not raised."""
    try:
        run_step(p, feats, init, form, work)
    except RuntimeError as e:
        p.synchronize()
        return None, {"failed": str(e)}
    p.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        it, nm, nfb, nnz = run_step(p, feats, init, form, work)
    p.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return dt, dict(irlba_iters=it, irlba_nmult=nm, knn_fallback_rows=nfb, snn_nnz=nnz)


def measured_traffic():
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    (scaled by the copy probe), from the `TRAFFIC <kernel> <form> <workload>
    <bytes>` lines of the committed profile summaries."""
    out = {}
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return out
    for fn in sorted(os.listdir(pdir)):
        if not fn.endswith(".md"):
            continue
        for line in open(os.path.join(pdir, fn)):
            f = line.split()
            if len(f) == 5 and f[0] == "TRAFFIC":
                out[(f[1], f[2], f[3])] = float(f[4])
    return out


def roofline(prof, form, steps, workload):
    if not prof:
        return None
    name, (launches, ms, by, fl) = max(prof.items(), key=lambda kv: kv[1][1])
    traffic = measured_traffic().get((name, form, workload))
    if fl > 0:
        achieved = fl / (ms * 1e-3) / 1e12
        # knn_cand_<form>[_kl48 or _kl64]: the arithmetic of that kernel
        kform = name.split("_")[2] if name.startswith("knn_cand_") else "fp32"
        peak = PEAK_TFLOPS[kform if kform in PEAK_TFLOPS else "fp32"]
        return {"kernel": name, "bound": "mfma", "achieved": round(achieved, 3), "peak": peak,
                "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": traffic,
                "launches_per_step": launches // steps,
                "avg_launch_ms": round(ms / launches, 4)}
    achieved = by / (ms * 1e-3) / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
            "launches_per_step": launches // steps, "avg_launch_ms": round(ms / launches, 4)}


def end_to_end(sev, p, feats, init, form, m):
    """One step with the counts handed over from host memory in Julia's
    layout (1-based Int64 colptr/rowval, Int32 values, as a ccall passes
    them) and the SNN graph copied back: the PCIe-inclusive rate."""
    mm, nn, colptr, rowval, vals = p.get_counts_julia()
    p.synchronize()
    t0 = time.perf_counter()
    p.load_counts_julia(mm, nn, colptr, rowval, vals)
    p.synchronize()
    t1 = time.perf_counter()
    run_step(p, feats, init, form)
    S = p.get_snn(m)
    t2 = time.perf_counter()
    del colptr, rowval, vals
    return {"cells_per_s": round(m / (t2 - t0), 2), "ms": round((t2 - t0) * 1e3, 1),
            "h2d_counts_ms": round((t1 - t0) * 1e3, 1), "snn_nnz": int(S.nnz), "form": form}


def fp64_errors(sev, p, npcs):
    """Singular values and right singular vectors of the pipeline's centered
    scaled matrix against an FP64 reference: the centered Gram matrix
    B'B - mu q' - q mu' + m mu mu' (q = column sums) accumulated in FP64 on
    the host from densified row blocks (BLAS) and eigendecomposed in FP64 --
    no Lanczos involved.  Returns a function of (s, V) giving the errors."""
    B, mu = p.get_scaled()
    m, n = B.shape
    Bc = B.tocsr()
    del B
    G = np.zeros((n, n))
    q = np.zeros(n)
    step = 65536
    for r0 in range(0, m, step):
        blk = Bc[r0:r0 + step].astype(np.float64).toarray()
        G += blk.T @ blk
        q += blk.sum(axis=0)
        del blk
    mu = mu.astype(np.float64)
    G = G - np.outer(mu, q) - np.outer(q, mu) + m * np.outer(mu, mu)
    w, V = np.linalg.eigh(G)
    sg = np.sqrt(np.maximum(w[::-1][:npcs], 0))
    Vg = V[:, ::-1][:, :npcs]

    sall = np.sqrt(np.maximum(w[::-1], 0))

    def flip(M):
        # svd_flip!(S; u_based_decision=false) (src/utils.jl:214-228): the
        # entry of largest magnitude of each column made positive
        i = np.argmax(np.abs(M), axis=0)
        return M * np.sign(M[i, np.arange(M.shape[1])])[None, :]

    def errors(s, Vf):
        rel = np.abs(s[:npcs] - sg) / sg
        qa, _ = np.linalg.qr(Vf[:, :npcs])
        cos = np.linalg.svd(qa.T @ Vg, compute_uv=False)
        # per component, for the ones 1% or more from both neighbours
        gap = np.minimum(np.r_[np.inf, (sall[:npcs - 1] - sall[1:npcs]) / sall[1:npcs]],
                         (sall[:npcs] - sall[1:npcs + 1]) / sall[:npcs])
        keep = np.nonzero(gap > 1e-2)[0]
        a, b = flip(np.asarray(Vf[:, :npcs], np.float64)), flip(Vg)
        ang = [float(np.arccos(min(1.0, abs(a[:, i] @ b[:, i]) /
                                   (np.linalg.norm(a[:, i]) * np.linalg.norm(b[:, i])))))
               for i in keep]
        agree = all(a[:, i] @ b[:, i] > 0 for i in keep)
        return {"sv_max_rel": float(rel.max()), "sv_mean_rel": float(rel.mean()),
                "subspace_angle_rad": float(np.arccos(np.clip(cos.min(), -1, 1))),
                "separated_components": [int(i) for i in keep],
                "svd_flip_angle_max_rad": max(ang) if ang else None,
                "svd_flip_signs_agree": bool(agree)}
    return errors


BATCH = dict(samples=64, genes=20000, cells=(500, 50000), density=(0.02, 0.08), features=2000,
             npcs=30, k=15, seed=1000, work=64)


def batch_leg(sev, steps, warmup, graph=True):
    """The batching form: 64 samples over one 20,000-gene set, cells
    log-uniform in [500, 50000], densities uniform in [2%, 8%], seed per
    sample = 1000 + s; per-sample normalise, 2000 features (each sample's
    top generator means), IRLBA 30 components with 64 Lanczos vectors per
    sample (BF16 operator values, FP32 left bases, FP64 arithmetic), kNN k = 15,
    SNN prune 1/15.
    The batch runs all samples in one
    stacked launch sequence (restart rounds replayed from HIP graphs); the
    comparison loops the single-sample pipeline over the samples."""
    b = BATCH
    rng = np.random.default_rng(b["seed"])
    S = b["samples"]
    lo, hi = b["cells"]
    cells = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), S))).astype(np.int64)
    dens = rng.uniform(b["density"][0], b["density"][1], S)
    tabs = [sev.nb_tables(b["genes"], seed=b["seed"] + s, density=float(dens[s]))
            for s in range(S)]
    feats = np.stack([sev.top_mean_features(t, b["features"]) for t in tabs])
    init = rng.standard_normal((S, b["features"]))
    M = int(cells.sum())

    pb = sev.Pipeline()
    pb.synth_batch(cells, tabs)
    _, _, nnz = pb.counts_info()

    def batched():
        pb.normalize("lognormalize", SF)
        pb.select_scale_batch(feats, SCALE_MAX)
        it, nm = pb.pca_batch(b["npcs"], init=init, tol=1e-5, graph=graph, work=b["work"])
        nfb = pb.neighbours(b["k"], include_self=True, precision="bf16")
        nz = pb.snn(PRUNE, np.float32)
        pb.synchronize()
        return it, nm, nfb, nz

    for _ in range(warmup):
        batched()
    pb.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        it, nm, nfb, nz = batched()
    dt_b = (time.perf_counter() - t0) / steps
    # phase breakdown from one more (untimed) step with every region timed
    pb.profile(True)
    batched()
    prof = pb.profile_read()
    pb.profile(False)
    for name, (l, ms, by, fl) in sorted(prof.items(), key=lambda kv: -kv[1][1])[:12]:
        print(f"#   batch {name:14s} {l:6d} launches/step {ms:9.2f} ms/step",
              file=sys.stderr, flush=True)
    pb.close()

    # the same samples through the single-sample pipeline, one after another
    ps = []
    for s in range(S):
        q = sev.Pipeline()
        q.synth_counts(int(cells[s]), tabs[s])
        ps.append(q)

    def looped(form):
        its = []
        for s, q in enumerate(ps):
            q.normalize("lognormalize", SF)
            q.select_scale(feats[s], SCALE_MAX)
            its.append(q.pca(b["npcs"], init=init[s], tol=1e-5, precision=form,
                             work=min(b["work"], int(cells[s]), b["features"]))[0])
            q.neighbours(b["k"], include_self=True, precision=form)
            q.snn(PRUNE, np.float32)
        for q in ps:
            q.synchronize()
        return its

    # one warm-up pass and one timed pass each (a pass is 64 pipelines)
    loops = {}
    for form in ("fp32", "bf16"):
        looped(form)
        t0 = time.perf_counter()
        looped(form)
        loops[form] = time.perf_counter() - t0
    for q in ps:
        q.close()
    out = {
        "workload": f"{S} samples x {b['genes']} genes, cells log-uniform in [{lo}, {hi}] "
                    f"(total {M}), density U({b['density'][0]}, {b['density'][1]}), "
                    f"{b['features']} features, {b['npcs']} PCs ({b['work']} Lanczos vectors, "
                    f"in the batch and in the loops), k={b['k']}, prune 1/15",
        "samples": S, "cells": M, "nnz": int(nnz),
        "ms_per_step": round(dt_b * 1e3, 3), "cells_per_s": round(M / dt_b, 2),
        "form": "bf16 operator values, bf16 kNN candidates, one stacked launch sequence, "
                + ("restart rounds replayed from HIP graphs" if graph else
                   "restart rounds launched directly"),
        "irlba_iters": [int(it.min()), int(np.median(it)), int(it.max())],
        "irlba_products": int(nm), "knn_fallback_rows": int(nfb), "snn_nnz": int(nz),
        "loop_plain_fp32": {"ms_per_step": round(loops["fp32"] * 1e3, 3),
                            "cells_per_s": round(M / loops["fp32"], 2)},
        "loop_bf16": {"ms_per_step": round(loops["bf16"] * 1e3, 3),
                      "cells_per_s": round(M / loops["bf16"], 2)},
        "speedup_vs_loop_plain_fp32": round(loops["fp32"] / dt_b, 3),
        "speedup_vs_loop_bf16": round(loops["bf16"] / dt_b, 3),
    }
    print(f"# batch: {out}", file=sys.stderr, flush=True)
    return out


def cpu_stages_full(orc, p, feats, m):
    """The oracle's linear stages on the whole workload, single-threaded:
    normalise, gather + scale, and one centered product pair (A v, A'w) --
    the unit of work an IRLBA product count multiplies.  Its kNN is
    O(cells^2) and is timed on the sample instead.  The counts are the
    device's, copied to the host."""
    X = p.get_counts()
    cp = np.ascontiguousarray(X.indptr, np.int64)
    rv = np.ascontiguousarray(X.indices, np.int32)
    cnt = np.ascontiguousarray(X.data, np.int32)
    nnz = len(rv)
    del X
    t0 = time.perf_counter()
    norm = orc.normalize(m, cp, rv, cnt, SF, True, np.float32)
    t_norm = time.perf_counter() - t0
    t0 = time.perf_counter()
    fcp, frow, fval = orc.gather(feats, cp, rv, norm)
    b, mu, _ = orc.scale(m, fcp, fval, np.float32(SCALE_MAX))
    t_scale = time.perf_counter() - t0
    del norm, cp, rv, cnt, fval
    x = np.random.default_rng(1).standard_normal(len(feats))
    t0 = time.perf_counter()
    y = orc.centered_mul(m, fcp, frow, b, mu, x)
    orc.centered_mul(m, fcp, frow, b, mu, y, trans=True)
    t_pair = time.perf_counter() - t0
    return {"cells": m, "nnz": nnz, "selected_nnz": int(fcp[-1]), "cores": 1, "kind": "port",
            "normalise_s": round(t_norm, 3), "gather_scale_s": round(t_scale, 3),
            "product_pair_s": round(t_pair, 3)}


def cpu_baseline(sev, orc, w, sample_cells, feats, threads=1):
    """The oracle's scalar C restatement of the same path on a bounded
    sample of the workload (same generator tables, first `sample_cells`
    cells' worth of draws, the timed step's features).  threads > 1 runs the
    kNN stage (the only threaded one) on that many threads; the other stages
    stay scalar, as the reference's Julia loops are."""
    t = sev.nb_tables(w["genes"], seed=0, density=w["density"], boost_frac=w["boost"])
    cp, rv, cnt = orc.synth_csc(sample_cells, t)
    init = np.random.default_rng(0).standard_normal(NFEAT)
    t0 = time.perf_counter()
    norm = orc.normalize(sample_cells, cp, rv, cnt, SF, True, np.float32)
    fcp, frow, fval = orc.gather(feats, cp, rv, norm)
    b, mu, _ = orc.scale(sample_cells, fcp, fval, np.float32(SCALE_MAX))
    rc, s, U, V, it, nm = orc.irlba_csc(sample_cells, fcp, frow, b, mu, NPCS, init, tol=1e-5)
    Z = np.asfortranarray((U * s).astype(np.float32))
    idx, _ = orc.knn(Z, K, nthreads=threads)
    orc.snn(idx, K, PRUNE, np.float32)
    dt = time.perf_counter() - t0
    return {"value": round(sample_cells / dt, 2), "unit": "cells/s", "cores": threads,
            "kind": "port",
            "sample": f"{sample_cells} cells of the same generator and genes, whole path, "
                      f"{dt:.1f} s (kNN is O(cells^2), so cells/s falls further at 1M cells)"
                      + (f"; kNN on {threads} threads, other stages scalar" if threads > 1 else
                         "")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="1m", choices=sorted(WORKLOADS))
    ap.add_argument("--forms", default=None, help="default: bf16,fp32 (fp8: FP8 candidates first; at d = 50 their "
                    "certificate passes no row, so every row also takes the BF16 pass)")
    ap.add_argument("--cpu-sample", type=int, default=8000)
    ap.add_argument("--maxit", type=int, default=None,
                    help="IRLBA restart limit (default 1000; 4m: 5000)")
    ap.add_argument("--work", type=int, default=DEFAULT_WORK,
                    help="Lanczos vectors kept by IRLBA (0: npcs + 7, the reference's choice)")
    ap.add_argument("--features", default="vst", choices=["vst", "top-mean"],
                    help="feature set of the timed step")
    ap.add_argument("--variant-steps", type=int, default=3,
                    help="steps of the secondary timings (reference work size, other feature "
                         "set); 0 skips them")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-e2e", action="store_true", help="skip the PCIe-inclusive step")
    ap.add_argument("--no-batch", action="store_true", help="skip the batching-form leg")
    ap.add_argument("--no-cluster", action="store_true",
                    help="skip cluster(SNN) (Louvain) on the last step's graph")
    ap.add_argument("--batch-direct", action="store_true",
                    help="launch the batch's restart rounds directly instead of from HIP graphs")
    ap.add_argument("--errors", action="store_true",
                    help="report singular-value and subspace errors against an FP64 Gram "
                         "eigendecomposition (host BLAS; adds about a minute at 1M cells)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the path is not sharded across "
                          "GPUs", "n_gpus": args.gpus}))
        sys.exit(2)

    sev = entry.load_package()
    sev.load()
    # known-size copy kernel: the device's achievable HBM rate, and the
    # calibration point of the byte counters in the profiles
    copy_gbs = sev.copy_probe()
    print(f"# HBM copy probe: {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)
    w = WORKLOADS[args.workload]
    global K, MAXIT, WORK
    WORK = args.work
    K = w.get("k", K)
    MAXIT = args.maxit if args.maxit is not None else w.get("maxit", 1000)
    if args.forms is None:
        args.forms = w.get("forms", "bf16,fp32")
    t = sev.nb_tables(w["genes"], seed=0, density=w["density"], boost_frac=w["boost"])
    p = sev.Pipeline()
    t0 = time.perf_counter()
    p.synth_counts(w["cells"], t)
    m, g, nnz = p.counts_info()
    gen_s = time.perf_counter() - t0
    top = sev.top_mean_features(t, NFEAT)
    init = np.random.default_rng(0).standard_normal(NFEAT)
    print(f"# generated {m} x {g} counts, nnz {nnz} ({nnz / (m * g):.4f}) in {gen_s:.1f} s",
          file=sys.stderr, flush=True)

    # find_variable_features(:vst) on the same counts, outside the timed
    # step (the feature list of the step is fixed in advance)
    if g > NFEAT:
        p.find_variable_features(NFEAT)
        p.synchronize()
        t0 = time.perf_counter()
        hvf = p.find_variable_features(NFEAT)
        hvg_ms = (time.perf_counter() - t0) * 1e3
        print(f"# find_variable_features(:vst, {NFEAT}): {hvg_ms:.1f} ms, "
              f"{len(set(hvf.tolist()) & set(top.tolist()))} of its genes among the top-mean "
              "ones", file=sys.stderr, flush=True)
    else:
        hvf, hvg_ms = top, 0.0  # every gene is a feature
    feats, other = (hvf, top) if args.features == "vst" else (top, hvf)
    fname = {"vst": "find_variable_features(:vst)", "top-mean": "largest generator mean"}

    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    res = {}
    errs = {}
    err_fn = None
    for form in forms:
        dt, info, prof, breakdown = time_form(p, feats, init, form, args.steps, args.warmup, True)
        res[form] = (dt, info, prof)
        if args.errors:
            try:
                if err_fn is None:
                    err_fn = fp64_errors(sev, p, NPCS)
                s_, _, V_, _ = p.get_pca()
                errs[form] = err_fn(s_, V_)
            except Exception as e:  # the report must not hide the timing
                errs[form] = {"error": str(e)[:200]}
        print(f"# {form}: {dt * 1e3:.1f} ms/step, {m / dt:.0f} cells/s, {info}", file=sys.stderr,
              flush=True)
        # phase breakdown of the untimed profiled step
        for name, (l, ms, by, fl) in sorted(breakdown.items(), key=lambda kv: -kv[1][1]):
            print(f"#   {name:14s} {l:6d} launches/step {ms:9.2f} ms/step"
                  f"  {by / (ms * 1e-3) / 1e9 if ms > 0 and by > 0 else 0:8.1f} GB/s"
                  f"  {fl / (ms * 1e-3) / 1e12 if ms > 0 and fl > 0 else 0:7.2f} TF/s",
                  file=sys.stderr, flush=True)
    head = forms[0]
    dt, info, prof = res[head]
    out = {
        "metric": "cells/s normalise\u2192PCA50\u2192kNN20\u2192SNN, BF16 vs FP32 form, 1M cells",
        "value": round(m / dt, 2),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": head,
        "data": "synthetic negative-binomial counts generated on the device",
        "config": {"workload": f"{args.workload}: {m} cells x {g} genes, nnz {nnz}, "
                               f"{NFEAT} features ({fname[args.features] if g > NFEAT else 'all genes'}), "
                               f"{NPCS} PCs (IRLBA, {WORK or NPCS + 7} Lanczos vectors), k={K}, "
                               f"prune 1/15, form {head}",
                   "cells": m, "genes": g, "nnz": nnz, "features": NFEAT,
                   "feature_set": args.features if g > NFEAT else "all", "npcs": NPCS,
                   "irlba_work": WORK or NPCS + 7, "k": K, "form": head},
        "roofline": dict(roofline(prof, head, args.steps, args.workload) or {}, copy_probe_gbs=round(copy_gbs, 1)),
        "forms": {f: {"cells_per_s": round(m / res[f][0], 2), "ms_per_step": round(res[f][0] * 1e3, 3),
                      **res[f][1]} for f in forms},
    }
    if "fp32" in res and "bf16" in res:
        out["bf16_speedup_vs_fp32"] = round(res["fp32"][0] / res["bf16"][0], 3)
    if "fp32" in res and "fp8" in res:
        out["fp8_speedup_vs_fp32"] = round(res["fp32"][0] / res["fp8"][0], 3)
    out["find_variable_features_ms"] = round(hvg_ms, 2)
    if errs:
        out["errors_vs_fp64"] = errs
    if not args.no_cluster:
        # Louvain (10 random starts) on the SNN graph of the last timed step,
        # downstream of the metric's path and outside its timing
        t0 = time.perf_counter()
        lab, q = p.cluster(m, return_modularity=True)
        cl_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, q1 = p.cluster(m, nrandomstarts=1, return_modularity=True)
        cl1_s = time.perf_counter() - t0
        out["cluster"] = {"seconds": round(cl_s, 3), "nrandomstarts": 10, "resolution": 0.8,
                          "clusters": int(lab.max()), "modularity": round(q, 6),
                          "seconds_1start": round(cl1_s, 3), "modularity_1start": round(q1, 6),
                          "graph": f"SNN of the {forms[-1]} step"}
        snn_host = None if args.no_cpu else p.get_snn(m)
        print(f"# cluster: {out['cluster']}", file=sys.stderr, flush=True)
    if not args.no_e2e:
        out["end_to_end"] = end_to_end(sev, p, feats, init, head, m)
        print(f"# end to end ({head}): {out['end_to_end']}", file=sys.stderr, flush=True)
    # secondary timings in the headline form: the reference's work size, and
    # the other feature set
    variants = {}
    if args.variant_steps > 0:
        if WORK != 0:
            vdt, vinfo = time_variant(p, feats, init, head, args.variant_steps, work=0)
            variants["reference_work"] = {"irlba_work": NPCS + 7, "maxit": MAXIT, **vinfo}
            if vdt is not None:
                variants["reference_work"].update(ms_per_step=round(vdt * 1e3, 3),
                                                  cells_per_s=round(m / vdt, 2))
            print(f"# {head} with work = npcs + 7: {variants['reference_work']}", file=sys.stderr,
                  flush=True)
        if g > NFEAT:
            vdt, vinfo = time_variant(p, other, init, head, args.variant_steps)
            oname = "top_mean_features" if args.features == "vst" else "vst_features"
            variants[oname] = dict(vinfo)
            if vdt is not None:
                variants[oname].update(ms_per_step=round(vdt * 1e3, 3),
                                       cells_per_s=round(m / vdt, 2))
            print(f"# {head} on the other feature set: {variants[oname]}", file=sys.stderr,
                  flush=True)
    if variants:
        out["variants"] = variants
    cpu_full = None
    if not args.no_cpu and args.workload in ("1m", "100k", "small"):
        cpu_full = cpu_stages_full(entry.load_oracle(), p, feats, m)
        print(f"# oracle stages on the whole workload: {cpu_full}", file=sys.stderr, flush=True)
    p.close()
    if not args.no_batch:
        out["batch"] = batch_leg(sev, max(1, min(args.steps, 2)), max(1, args.warmup),
                                 graph=not args.batch_direct)
    if not args.no_cpu:
        orc = entry.load_oracle()
        out["cpu_baseline"] = cpu_baseline(sev, orc, w, args.cpu_sample, feats)
        if cpu_full:
            out["cpu_baseline"]["stages_full_workload"] = cpu_full
        nthr = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        if nthr > 1:
            out["cpu_baseline_all_cores"] = cpu_baseline(sev, orc, w, args.cpu_sample, feats,
                                                         nthr)
        if not args.no_cluster:
            # the oracle's sequential Louvain (one start, one core) on the same graph
            S = snn_host.tocsc()
            S.sort_indices()
            t0 = time.perf_counter()
            _, qo = orc.louvain(m, S.indptr, S.indices, S.data.astype(np.float64), nstarts=1)
            out["cluster"]["cpu_port_1start"] = {"seconds": round(time.perf_counter() - t0, 3),
                                                 "modularity": round(qo, 6), "cores": 1,
                                                 "kind": "port"}
            print(f"# cluster, oracle one start: {out['cluster']['cpu_port_1start']}",
                  file=sys.stderr, flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
