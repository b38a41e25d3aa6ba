"""The reference's own whole-fit quality protocol for its CI smoke gates, so
the oracle and the GPU fit can be judged against the numbers the reference
published for itself.

Restated (not imported) from the reference's benchmark scripts:
  datasets         benchmarks/ci_ann_smoke.py:26-40 (and ci_consistency_smoke.py:26-40):
                   sklearn's digits / breast_cancer, StandardScaler on float32
  CSV round trip   benchmarks/compare_real_impls_fair.py:130-131 (np.savetxt
                   "%.8f"; the Rust binaries parse the text back into f32)
  shared exact kNN compare_real_impls_fair.py:320-339 (sklearn brute force,
                   self dropped: k columns of non-self neighbours)
  run settings     compare_real_impls_fair.py:59-62, 341-375, 414-445:
                   k = 15, 2-D, 200 epochs, random init, seed 42; the ann-e2e
                   gate runs fit_csv in auto mode (exact below 16,384 rows),
                   the consistency gate bench_fit_csv on the shared kNN
  consistency      compare_real_impls_fair.py:457-541: a default_rng(seed)
                   sub-sample of sample_cap rows, trustworthiness@15 and
                   knn_recall_vs_original, whose embedding-side kNN
                   (knn_indices) asks sklearn for k + 1 neighbours of the
                   fitted points (self already excluded) and drops the first,
                   i.e. it compares ranks 2..k+1 of the embedding; that quirk
                   is kept so the figures are comparable
The published figures (benchmarks/ann-e2e-smoke.eval.json:46-55, 91-100 and
benchmarks/consistency-smoke.eval.json) are single seed-42 runs of the Rust
CPU path; tests/golden/reference_smoke.json holds them with the oracle's
figures over several seeds.
"""
from __future__ import annotations

import numpy as np

K = 15

# gate -> (sample_cap, uses the shared precomputed kNN)
GATES = {"ann_e2e": (1200, False), "consistency": (1500, True)}

# benchmarks/ann-e2e-smoke.eval.json:46-55, 91-100 and
# benchmarks/consistency-smoke.eval.json (rust_umap entries)
REFERENCE = {
    "ann_e2e": {
        "digits": {"trust": 0.9660143018974795, "recall": 0.46688888888888785},
        "breast_cancer": {"trust": 0.9143041687857154, "recall": 0.4063268892794386},
    },
    "consistency": {
        "digits": {"trust": 0.9700581960430301, "recall": 0.4541333333333326},
        "breast_cancer": {"trust": 0.9225177088094487, "recall": 0.4123022847100187},
    },
}


def _csv_round_trip(a: np.ndarray) -> np.ndarray:
    """np.savetxt(fmt='%.8f') then a parse back to f32."""
    txt = np.char.mod("%.8f", np.asarray(a, dtype=np.float64))
    return txt.astype(np.float64).astype(np.float32)


def load_dataset(name: str) -> np.ndarray:
    from sklearn.datasets import load_breast_cancer, load_digits
    from sklearn.preprocessing import StandardScaler

    raw = {"digits": load_digits, "breast_cancer": load_breast_cancer}[name]().data
    x = StandardScaler().fit_transform(raw.astype(np.float32)).astype(np.float32)
    return _csv_round_trip(x)


def shared_exact_knn(x: np.ndarray, k: int = K):
    """compute_shared_exact_knn: k non-self neighbours (int64 indices, f32
    distances after the CSV round trip)."""
    from sklearn.neighbors import NearestNeighbors

    nn = NearestNeighbors(n_neighbors=k + 1, algorithm="brute", metric="euclidean", n_jobs=1)
    nn.fit(x)
    dist, idx = nn.kneighbors(x)
    return (np.ascontiguousarray(idx[:, 1:k + 1].astype(np.int64)),
            _csv_round_trip(dist[:, 1:k + 1].astype(np.float32)))


def _knn_indices(emb: np.ndarray, k: int) -> np.ndarray:
    from sklearn.neighbors import NearestNeighbors

    nn = NearestNeighbors(n_neighbors=k + 1, metric="euclidean")
    nn.fit(emb)
    return nn.kneighbors(return_distance=False)[:, 1:]


def _recall(orig: np.ndarray, emb: np.ndarray, k: int) -> float:
    e = _knn_indices(emb, k)
    return float(np.mean([len(set(a.tolist()) & set(b.tolist())) / k for a, b in zip(orig, e)]))


def consistency(x: np.ndarray, emb: np.ndarray, orig_knn_idx: np.ndarray, sample_cap: int,
                seed: int = 42, k: int = K) -> tuple[float, float]:
    """(trustworthiness@k, original-kNN recall@k) as compute_consistency."""
    from sklearn.manifold import trustworthiness

    n = x.shape[0]
    if n > sample_cap:
        sel = np.random.default_rng(seed).choice(n, size=sample_cap, replace=False)
        xs, es = x[sel], emb[sel]
        orig = _knn_indices(xs, k)
    else:
        xs, es = x, emb
        orig = orig_knn_idx
    return float(trustworthiness(xs, es, n_neighbors=k)), _recall(orig, es, k)
