"""Writes the golden fixtures of this directory from the oracle (the C++
restatement of the reference's CPU path under oracle/).

    python tests/golden/make_golden.py          # all fixtures (about 15 minutes)
    python tests/golden/make_golden.py c4_sub50k_k30_200   # one quality section

Fixtures (all small):
  c1_stages.npz  config-1 (1000 x 32 i.i.d. N(0,1), seed 42) stage outputs:
                 kNN (k=15) indices/distances, sigma/rho, the pruned fuzzy
                 graph for 200 epochs (heads, tails, weights), and the per-edge
                 positive / negative sample counts of 200 epochs
  inputs.json    SHA-256 of the synthetic inputs of configs 1-3 and of the
                 100,000-row fit set of config 5
  reference_smoke.json
                 the oracle under the reference's CI smoke protocol (digits and
                 breast_cancer, k = 15, 200 epochs, random init; both gates)
                 over ten seeds, beside the reference's published seed-42
                 figures for its Rust CPU path
  quality.json   the oracle's trustworthiness@15 and kNN preservation at
                 config 1, config 2 (densMAP, 500 epochs) and a 50,000-row
                 draw of config 4's mixture (datasets.c4(n=50000), k = 30),
                 configs 3 (euclidean and cosine) and 5 (with the embedding's
                 extent) over fixed seeds (the statistical
                 parity targets of the GPU fit tests)

The reference (Rust) is not run: it holds no array fixtures of its own for
this path, and its toolchain is not available here.  Each fixture is checked
independently by tests/test_golden.py (FP64 brute force, closed forms and
invariants), so a fixture that the oracle wrote wrongly fails there.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from umapers_amd import datasets  # noqa: E402
from umapers_amd.diagnostics import knn_preservation, trustworthiness  # noqa: E402

C1_SEEDS = list(range(40))
C2_SEEDS = [42, 7, 11, 23, 5, 9, 13, 17, 19, 29, 31, 37, 41, 43, 47, 53,
            59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113, 127, 131]
C4_SEEDS = [42, 7, 11, 23]


def stages() -> None:
    x = datasets.c1()
    idx, dist = O.knn(x, 15)
    sig, rho = O.smooth_knn_dist(dist, 15.0)
    h, t, w = O.fuzzy_graph(idx, dist, sig, rho, 1.0, 200)
    pos, neg, _, _ = O.schedule_counts(w, 200, 5)
    np.savez_compressed(os.path.join(HERE, "c1_stages.npz"), knn_idx=idx.astype(np.int32),
                        knn_dist=dist, sigma=sig, rho=rho, heads=h.astype(np.int32),
                        tails=t.astype(np.int32), weights=w, pos_counts=pos.astype(np.int32),
                        neg_counts=neg.astype(np.int32))


def inputs() -> None:
    out = {"c1": datasets.sha256(datasets.c1()), "c2": datasets.sha256(datasets.c2()),
           "c3": datasets.sha256(datasets.c3()),
           "c5_fit": datasets.sha256(datasets.c5(n_fit=100_000, n_query=1)[0])}
    with open(os.path.join(HERE, "inputs.json"), "w") as f:
        json.dump(out, f, indent=1)


def _fit_quality(x, seeds, **kw) -> dict:
    t, k = [], []
    for s in seeds:
        y = O.fit(x, O.OracleParams(random_seed=s, **kw)).embedding
        t.append(trustworthiness(x, y))
        k.append(knn_preservation(x, y))
        print(kw, "seed", s, t[-1], flush=True)
    return {"seeds": list(seeds), "trust": t, "knn_preservation": k}


def _fit_quality_shared_knn(x, seeds, metric: str, **kw) -> dict:
    """As _fit_quality, with the exact kNN computed once and handed to every
    seed's fit (the kNN does not depend on the seed; fit_transform_with_knn
    runs the same stages after it)."""
    idx, dist = O.knn(x, kw.get("n_neighbors", 15), metric)
    t, k = [], []
    for s in seeds:
        y = O.fit(x, O.OracleParams(random_seed=s, metric=metric, **kw), knn_idx=idx,
                  knn_dist=dist).embedding
        t.append(trustworthiness(x, y))
        k.append(knn_preservation(x, y))
        print(metric, kw, "seed", s, t[-1], k[-1], flush=True)
    return {"seeds": list(seeds), "trust": t, "knn_preservation": k}


C3_SEEDS = [42, 7, 11]
C5_SEEDS = [42, 7, 11]


def _radius_max(y) -> float:
    """Largest distance of an embedded point from the embedding's median."""
    return float(np.linalg.norm(y - np.median(y, 0), axis=1).max())


def _radius_q(y, q) -> float:
    """The q-th percentile of the embedded points' distances from the median."""
    return float(np.percentile(np.linalg.norm(y - np.median(y, 0), axis=1), q))


def _fit_quality_c5(seeds) -> dict:
    """Config 5's 100,000 fit rows: quality and the embedding's extent (its
    high-dimensional kNN graph has hub rows of about a thousand entries)."""
    x = datasets.c5(n_fit=100_000, n_query=1)[0]
    idx, dist = O.knn(x, 15)
    t, k, r, r50, r99 = [], [], [], [], []
    for s in seeds:
        y = O.fit(x, O.OracleParams(random_seed=s, n_epochs=200), knn_idx=idx,
                  knn_dist=dist).embedding
        t.append(trustworthiness(x, y))
        k.append(knn_preservation(x, y))
        r.append(_radius_max(y))
        r50.append(_radius_q(y, 50))
        r99.append(_radius_q(y, 99))
        print("c5 seed", s, t[-1], k[-1], r[-1], r50[-1], r99[-1], flush=True)
    return {"seeds": list(seeds), "trust": t, "knn_preservation": k, "radius_max": r,
            "radius_p50": r50, "radius_p99": r99}

QUALITY = {
    "c1_plain_200": lambda: _fit_quality(datasets.c1(), C1_SEEDS, n_epochs=200),
    "c3_euclidean_200": lambda: _fit_quality_shared_knn(datasets.c3(), C3_SEEDS, "euclidean",
                                                        n_epochs=200),
    "c3_cosine_200": lambda: _fit_quality_shared_knn(datasets.c3(cosine=True), C3_SEEDS,
                                                     "cosine", n_epochs=200),
    "c5_fit_200": lambda: _fit_quality_c5(C5_SEEDS),
    "c2_densmap_500": lambda: _fit_quality(datasets.c2(), C2_SEEDS, n_epochs=500, densmap=True),
    "c4_sub50k_k30_200": lambda: _fit_quality(datasets.c4(n=50_000), C4_SEEDS, n_epochs=200,
                                              n_neighbors=30),
}


SMOKE_SEEDS = [42, 0, 1, 2, 3, 4, 5, 6, 7, 8]


def reference_smoke() -> None:
    """The oracle under the reference's own CI smoke protocol
    (tests/smoke_protocol.py), next to the figures the reference published
    for its Rust CPU path: reference_smoke.json."""
    from tests import smoke_protocol as P

    out = {"seeds": SMOKE_SEEDS, "k": P.K, "n_epochs": 200, "init": "random", "gates": {}}
    for gate, (cap, pre) in P.GATES.items():
        out["gates"][gate] = {"sample_cap": cap, "precomputed_knn": pre, "datasets": {}}
        for name in ("digits", "breast_cancer"):
            x = P.load_dataset(name)
            idx, dist = P.shared_exact_knn(x)
            ts, rs = [], []
            for s in SMOKE_SEEDS:
                prm = O.OracleParams(random_seed=s, n_epochs=200, init="random")
                y = (O.fit(x, prm, knn_idx=idx, knn_dist=dist) if pre else O.fit(x, prm)).embedding
                t, r = P.consistency(x, y, idx, cap)
                ts.append(t)
                rs.append(r)
            out["gates"][gate]["datasets"][name] = {
                "reference": P.REFERENCE[gate][name],
                "oracle_trust": ts, "oracle_recall": rs,
                "oracle_trust_mean": float(np.mean(ts)), "oracle_recall_mean": float(np.mean(rs))}
            print(gate, name, np.mean(ts), np.mean(rs), P.REFERENCE[gate][name], flush=True)
    with open(os.path.join(HERE, "reference_smoke.json"), "w") as f:
        json.dump(out, f, indent=1)


def quality(only=None) -> None:
    path = os.path.join(HERE, "quality.json")
    res = {}
    if os.path.exists(path):
        with open(path) as f:
            res = json.load(f)
    for key, make in QUALITY.items():
        if only is None or key in only:
            res[key] = make()
            with open(path, "w") as f:  # after each section, so a long run keeps its progress
                json.dump(res, f, indent=1)


if __name__ == "__main__":
    if sys.argv[1:] == ["reference_smoke"]:
        reference_smoke()
    elif len(sys.argv) > 1:  # named quality sections only, e.g. c4_sub50k_k30_200
        quality(set(sys.argv[1:]))
    else:
        stages()
        inputs()
        reference_smoke()
        quality()
