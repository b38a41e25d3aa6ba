"""Exact kNN on the GPU against the oracle (bit-exact) and FP64 (tie-tolerant)."""
import numpy as np
import pytest

from oracle.knn_fp64 import knn_fp64
from umapers_amd import datasets
from umapers_amd.diagnostics import recall_at_k

from . import capi

pytestmark = pytest.mark.gpu

SHAPES = [(300, 8, 5), (500, 32, 15), (777, 37, 15), (1000, 64, 15), (900, 100, 15),
          (1200, 128, 15), (1500, 192, 15), (600, 256, 15), (800, 64, 30)]


@pytest.mark.parametrize("n,d,k", SHAPES)
@pytest.mark.parametrize("metric", [0, 1])
def test_fp32_knn_bit_exact_vs_oracle(gpu, oracle, rng, n, d, k, metric):
    x = rng.normal(size=(n, d)).astype(np.float32)
    idx, dist = capi.knn(x, k, metric, 0)
    oi, od = oracle.knn(x, k, "cosine" if metric else "euclidean")
    assert np.array_equal(idx, oi)
    assert np.array_equal(dist.view(np.int32), od.view(np.int32))


@pytest.mark.parametrize("metric", [0, 1])
def test_bf16_rerank_knn_recall_and_exact_distances(gpu, oracle, metric):
    x = datasets.gaussian_mixture(3000, 128, 8, 3.0, 5, (0.5, 1.5))
    idx, dist = capi.knn(x, 15, metric, 1, 2)
    oi, od = oracle.knn(x, 15, "cosine" if metric else "euclidean")
    assert recall_at_k(idx, oi) >= 0.99
    # every returned distance is the exact f32 distance of that pair
    same = idx == oi
    assert np.array_equal(dist[same].view(np.int32), od[same].view(np.int32))


def test_fp32_knn_vs_fp64_ties_only(gpu):
    x = datasets.gaussian_mixture(2000, 128, 4, 3.0, 9, 1.0)
    idx, dist = capi.knn(x, 15)
    i64, d64 = knn_fp64(x, None, 15)
    bad = 0
    x64 = x.astype(np.float64)
    for r in np.where(np.any(idx != i64, axis=1))[0]:
        for j in set(idx[r]) ^ set(i64[r]):
            dj = np.sqrt(((x64[r] - x64[j]) ** 2).sum())
            if abs(dj - d64[r, -1]) > 1e-6 * d64[r, -1]:
                bad += 1
    assert bad == 0
    assert np.all(np.abs(dist - d64) <= 1e-5 * d64 + 1e-6)


def test_knn_to_reference_matches_oracle(gpu, oracle, rng):
    r = rng.normal(size=(1500, 96)).astype(np.float32)
    q = rng.normal(size=(333, 96)).astype(np.float32)
    for metric in (0, 1):
        idx, dist = capi.knn_to_reference(q, r, 15, metric)
        oi, od = oracle.knn_to_reference(q, r, 15, "cosine" if metric else "euclidean")
        assert np.array_equal(idx, oi)
        assert np.array_equal(dist, od)


def test_duplicate_rows_and_ties(gpu, oracle):
    base = np.random.default_rng(4).normal(size=(50, 16)).astype(np.float32)
    x = np.concatenate([base, base, base[:10]], axis=0)  # exact duplicates
    idx, dist = capi.knn(x, 15)
    oi, od = oracle.knn(x, 15)
    assert np.array_equal(idx, oi) and np.array_equal(dist, od)
    # a duplicated row's lower-index twin comes first: include_self is False
    assert not capi.include_self(idx, dist)
    # tie-break KAT (sparse.rs:1070-1122)
    ref = np.array([[1.0, 0.0], [-1.0, 0.0], [0.0, 1.0], [0.0, -1.0]], np.float32)
    qi, qd = capi.knn_to_reference(np.zeros((1, 2), np.float32), ref, 3)
    assert qi[0].tolist() == [0, 1, 2] and np.all(qd == 1.0)


def test_smallest_sizes(gpu, oracle):
    x = np.random.default_rng(5).normal(size=(16, 3)).astype(np.float32)
    idx, dist = capi.knn(x, 15)
    oi, od = oracle.knn(x, 15)
    assert np.array_equal(idx, oi) and np.array_equal(dist, od)


def test_cosine_zero_rows(gpu, oracle, rng):
    x = rng.normal(size=(200, 40)).astype(np.float32)
    x[[3, 17]] = 0.0
    idx, dist = capi.knn(x, 10, 1)
    oi, od = oracle.knn(x, 10, "cosine")
    # zero rows: distance 1 to every non-zero row, 0 to each other
    assert np.array_equal(dist, od)
    assert recall_at_k(idx, oi) > 0.99


def test_knn_to_reference_many_queries(gpu, oracle, rng):
    # 65536 queries fill one block per CU without splitting the references;
    # the small shapes above exercise the split-and-merge path
    r = rng.normal(size=(2000, 32)).astype(np.float32)
    q = rng.normal(size=(65536, 32)).astype(np.float32)
    for metric in (0, 1):
        idx, dist = capi.knn_to_reference(q, r, 15, metric)
        oi, od = oracle.knn_to_reference(q, r, 15, "cosine" if metric else "euclidean")
        assert np.array_equal(idx, oi)
        assert np.array_equal(dist, od)


# 17 .. 32 neighbours: four top-16 lists per query (two interleaved splits of
# the reference rows, two halves each), with the queries whose lists may have
# dropped one of their 32 best scores redone by the 32-entry kernel.
K32_SHAPES = [(2000, 32, 17), (1500, 100, 20), (3000, 64, 24), (2500, 128, 32), (40, 8, 30), (20, 5, 17)]


@pytest.mark.parametrize("n,d,k", K32_SHAPES)
@pytest.mark.parametrize("metric", [0, 1])
def test_17_to_32_neighbours_bit_exact(gpu, oracle, rng, n, d, k, metric):
    x = rng.normal(size=(n, d)).astype(np.float32)
    idx, dist = capi.knn(x, k, metric)
    oi, od = oracle.knn(x, k, "cosine" if metric else "euclidean")
    assert np.array_equal(idx, oi) and np.array_equal(dist, od)


def test_17_to_32_neighbours_sorted_and_duplicated_rows(gpu, oracle):
    # rows sorted by cluster (a query's neighbours sit together in the row
    # order), and 40 copies of each of 60 rows (every query has 39 ties at
    # distance 0, so every list is full of its 32 best: all redone)
    x, lab = datasets.gaussian_mixture(4000, 48, 20, 3.0, 3, (0.5, 1.5), True)
    x = np.ascontiguousarray(x[np.argsort(lab, kind="stable")])
    for xs in (x, np.repeat(np.random.default_rng(8).normal(size=(60, 24)).astype(np.float32), 40, axis=0)):
        for metric in (0, 1):
            idx, dist = capi.knn(xs, 30, metric)
            oi, od = oracle.knn(xs, 30, "cosine" if metric else "euclidean")
            assert np.array_equal(idx, oi) and np.array_equal(dist, od)
        # BF16 candidates with the exact re-rank
        idx, dist = capi.knn(xs, 30, 0, 1, 2)
        oi, od = oracle.knn(xs, 30)
        assert recall_at_k(idx, oi) >= 0.99


@pytest.mark.parametrize("metric", [0, 1])
def test_17_to_32_neighbours_bf16_recall(gpu, oracle, metric):
    x = datasets.gaussian_mixture(5000, 64, 8, 3.0, 5, (0.5, 1.5))
    idx, dist = capi.knn(x, 30, metric, 1, 2)
    oi, od = oracle.knn(x, 30, "cosine" if metric else "euclidean")
    assert recall_at_k(idx, oi) >= 0.99
    same = idx == oi
    assert np.array_equal(dist[same].view(np.int32), od[same].view(np.int32))


# Rows wider than 256 features take the GEMM score-block path with a
# selection kernel; the re-rank makes FP32 results equal to the oracle's.
WIDE = [(700, 300, 15), (1100, 784, 15), (400, 520, 30), (257, 257, 5), (20, 300, 5)]


@pytest.mark.parametrize("n,d,k", WIDE)
@pytest.mark.parametrize("metric", [0, 1])
def test_wide_rows_fp32_bit_exact_vs_oracle(gpu, oracle, rng, n, d, k, metric):
    x = rng.normal(size=(n, d)).astype(np.float32)
    idx, dist = capi.knn(x, k, metric, 0)
    oi, od = oracle.knn(x, k, "cosine" if metric else "euclidean")
    assert np.array_equal(idx, oi)
    assert np.array_equal(dist.view(np.int32), od.view(np.int32))


@pytest.mark.parametrize("metric", [0, 1])
def test_wide_rows_bf16_recall(gpu, oracle, metric):
    x = datasets.gaussian_mixture(2500, 400, 8, 3.0, 7, (0.5, 1.5))
    idx, dist = capi.knn(x, 15, metric, 1, 2)
    oi, od = oracle.knn(x, 15, "cosine" if metric else "euclidean")
    assert recall_at_k(idx, oi) >= 0.99
    same = idx == oi
    assert np.array_equal(dist[same].view(np.int32), od[same].view(np.int32))


# BF16 rows of 257 .. 896 features run the candidate kernel itself (query
# fragments in registers, one-tile groups in LDS); 900 features, and 32-entry
# lists past 384, stay on the score blocks.
WIDE_BF16 = [(1100, 784, 15), (400, 520, 30), (300, 380, 24), (257, 257, 5), (20, 300, 5), (1500, 900, 15),
             (600, 700, 20)]


@pytest.mark.parametrize("n,d,k", WIDE_BF16)
@pytest.mark.parametrize("metric", [0, 1])
def test_wide_rows_bf16_candidates(gpu, oracle, n, d, k, metric):
    x = datasets.gaussian_mixture(n, d, 8, 3.0, 7, (0.5, 1.5))
    idx, dist = capi.knn(x, k, metric, 1, 2)
    oi, od = oracle.knn(x, k, "cosine" if metric else "euclidean")
    assert recall_at_k(idx, oi) >= 0.99
    same = idx == oi
    assert np.array_equal(dist[same].view(np.int32), od[same].view(np.int32))


def test_wide_rows_bf16_duplicates(gpu, oracle, rng):
    # 300 copies of one 784-feature row: identical rows score identically in
    # the candidate kernel, so the ties resolve by index as the oracle's do
    base = rng.normal(size=(1, 784)).astype(np.float32)
    x = np.concatenate([np.repeat(base, 300, axis=0),
                        rng.normal(size=(400, 784)).astype(np.float32)])
    idx, dist = capi.knn(x, 10, 0, 1, 2)
    oi, od = oracle.knn(x, 10)
    assert np.array_equal(idx[:300], oi[:300]) and np.array_equal(dist[:300], od[:300])
    assert recall_at_k(idx[300:], oi[300:]) >= 0.99


def test_wide_rows_bf16_to_reference(gpu, oracle):
    # separate query rows (the transform's kNN) on the wide candidate path
    r = datasets.gaussian_mixture(1200, 600, 8, 3.0, 3, (0.5, 1.5))
    q = datasets.mixture_queries(150, 600, 8, 3.0, 3, (0.5, 1.5), 4)
    for metric in (0, 1):
        idx, dist = capi.knn_to_reference(q, r, 15, metric, 1)
        oi, od = oracle.knn_to_reference(q, r, 15, "cosine" if metric else "euclidean")
        assert recall_at_k(idx, oi) >= 0.99
        same = idx == oi
        assert np.array_equal(dist[same].view(np.int32), od[same].view(np.int32))


def test_wide_rows_many_duplicates(gpu, oracle, rng):
    # 1500 copies of one row: their score rows hold more than a thousand
    # equal values at the top, which the selection resolves by bisection
    base = rng.normal(size=(1, 320)).astype(np.float32)
    x = np.concatenate([np.repeat(base, 1500, axis=0),
                        rng.normal(size=(600, 320)).astype(np.float32)])
    idx, dist = capi.knn(x, 10, 0, 0)
    oi, od = oracle.knn(x, 10)
    assert np.array_equal(idx, oi)
    assert np.array_equal(dist, od)


def test_wide_rows_to_reference(gpu, oracle, rng):
    r = rng.normal(size=(900, 600)).astype(np.float32)
    q = rng.normal(size=(130, 600)).astype(np.float32)
    for metric in (0, 1):
        idx, dist = capi.knn_to_reference(q, r, 15, metric)
        oi, od = oracle.knn_to_reference(q, r, 15, "cosine" if metric else "euclidean")
        assert np.array_equal(idx, oi)
        assert np.array_equal(dist, od)


def test_wide_rows_fit_parity(gpu, oracle):
    from umapers_amd import Umap
    from umapers_amd.diagnostics import trustworthiness
    x = datasets.gaussian_mixture(3000, 784, 8, 3.0, 13, (0.5, 1.5))
    g = Umap(n_epochs=200, random_seed=5, ann_mode="exact").fit_transform(x)
    fit = oracle.fit(x, oracle.OracleParams(n_epochs=200, random_seed=5))
    assert abs(trustworthiness(x, g) - trustworthiness(x, fit.embedding)) <= 0.01


@pytest.mark.parametrize("metric", [0, 1])
def test_wide_rows_several_score_blocks(gpu, oracle, rng, monkeypatch, metric):
    # a score block of 192 query rows (the test hook lowers the block budget),
    # so 3000 rows run 15 full blocks and a partial one; FP32 stays bit-exact
    monkeypatch.setenv("UMAPX_TEST_WIDE_SCORE_BYTES", str(192 * 3000 * 4))
    x = rng.normal(size=(3000, 300)).astype(np.float32)
    idx, dist = capi.knn(x, 15, metric, 0)
    oi, od = oracle.knn(x, 15, "cosine" if metric else "euclidean")
    assert np.array_equal(idx, oi)
    assert np.array_equal(dist.view(np.int32), od.view(np.int32))
    q = rng.normal(size=(700, 300)).astype(np.float32)
    idx, dist = capi.knn_to_reference(q, x, 15, metric, 1)
    oi, od = oracle.knn_to_reference(q, x, 15, "cosine" if metric else "euclidean")
    assert recall_at_k(idx, oi) >= 0.99


BIG_K = [(700, 16, 33), (900, 70, 48), (600, 128, 64), (500, 300, 40)]


@pytest.mark.parametrize("n,d,k", BIG_K)
@pytest.mark.parametrize("metric", [0, 1])
def test_more_than_32_neighbours_bit_exact(gpu, oracle, rng, n, d, k, metric):
    # n_neighbors 33-64: the exact distances on the vector ALUs (euclidean
    # ranked on the squared distance, its eight lane sums for d >= 64;
    # cosine from the dot product and the rows' norms), bit-exact
    x = rng.normal(size=(n, d)).astype(np.float32)
    x[5] = x[9]  # a duplicate pair: a tie resolved by index
    name = ["euclidean", "cosine"][metric]
    gi, gd = capi.knn(x, k, metric, 0, 2)
    oi, od = oracle.knn(x, k, name)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)
    q = rng.normal(size=(37, d)).astype(np.float32)
    gi, gd = capi.knn_to_reference(q, x, k, metric)
    oi, od = oracle.knn_to_reference(q, x, k, name)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)


@pytest.mark.parametrize("metric", ["manhattan", "chebyshev", "correlation", "canberra",
                                    "braycurtis"])
def test_more_than_32_neighbours_other_metrics(gpu, oracle, rng, metric):
    x = rng.normal(size=(800, 72)).astype(np.float32)
    gi, gd = capi.knn_metric(x, x, 50, metric)
    oi, od = oracle.knn_metric(x, x, 50, metric)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)
