"""Pins the CPU oracle against the reference's own known answers and against
closed forms / textbook routines (CPU only)."""
import struct

import numpy as np
import pytest

from oracle.knn_fp64 import knn_fp64


def f32_bits(v: float) -> int:
    return struct.unpack("<I", struct.pack("<f", v))[0]


# rust_umap/src/lib.rs:5781-5795 optimized_curve_search_preserves_reference_results
AB_KAT = [
    (1.0, 0.1, 0x3FC9D9DF, 0x3F65217D),
    (1.0, 0.0, 0x3FF762F1, 0x3F4A5EA8),
    (2.0, 0.5, 0x3E848A83, 0x3F875BF1),
    (0.75, 0.2, 0x3FFFFFFE, 0x3F89AE8F),
    (3.0, 1.0, 0x3D95A936, 0x3F930F0C),
]


@pytest.mark.parametrize("spread,min_dist,ea,eb", AB_KAT)
def test_find_ab_params_bit_exact(oracle, spread, min_dist, ea, eb):
    a, b = oracle.find_ab_params(spread, min_dist)
    assert f32_bits(a) == ea and f32_bits(b) == eb


def test_smooth_knn_self_and_no_self_agree(oracle):
    # lib.rs:6975-7000 smooth_knn_dist_handles_rows_without_self_neighbor
    with_self = np.array([[0.0, 0.2, 0.4, 0.8], [0.0, 0.1, 0.5, 0.9]], np.float32)
    without = np.array([[0.2, 0.4, 0.8], [0.1, 0.5, 0.9]], np.float32)
    s1, r1 = oracle.smooth_knn_dist(with_self, 3.0, 1.0, include_self=True)
    s2, r2 = oracle.smooth_knn_dist(without, 3.0, 1.0, include_self=False)
    assert np.all(np.abs(s1 - s2) <= 1e-5) and np.all(np.abs(r1 - r2) <= 1e-6)
    # rho is the first positive distance
    assert np.allclose(r2, [0.2, 0.1])


def test_smooth_knn_zero_rows_agree(oracle):
    # lib.rs:7000-7020
    with_self = np.array([[0.0, 0.0, 0.0, 0.0], [0.0, 1.0, 2.0, 3.0]], np.float32)
    without = np.array([[0.0, 0.0, 0.0], [1.0, 2.0, 3.0]], np.float32)
    s1, r1 = oracle.smooth_knn_dist(with_self, 3.0, 1.0, include_self=True)
    s2, r2 = oracle.smooth_knn_dist(without, 3.0, 1.0, include_self=False)
    assert np.all(np.abs(s1 - s2) <= 1e-6) and np.all(np.abs(r1 - r2) <= 1e-6)


def test_smooth_knn_psum_hits_log2k(oracle, rng):
    # invariant: sum_j exp(-(d_j - rho)/sigma) == log2(k) within 1e-5 unless
    # sigma was floored or the 64 steps ran out
    d = np.sort(rng.uniform(0.5, 3.0, size=(200, 15)).astype(np.float32), axis=1)
    d[:, 0] = 0.0
    sig, rho = oracle.smooth_knn_dist(d, 15.0, 1.0, include_self=True)
    act = d[:, 1:].astype(np.float64)
    psum = np.where(act - rho[:, None] > 0, np.exp(-(act - rho[:, None]) / sig[:, None]), 1.0).sum(1)
    assert np.all(np.abs(psum - np.log2(15.0)) < 2e-5)
    assert np.array_equal(rho, d[:, 1])


def test_smooth_knn_fp64_sums_variant(oracle, rng):
    """The f64-psum form of the bisection (the GPU path's default): the same
    invariant with the sums in f64, the self / no-self equivalence, a
    closed-form row, and agreement with the f32 form (both solve the same
    equation, so they differ by the stop tolerance only)."""
    d = np.sort(rng.uniform(0.5, 3.0, size=(300, 15)).astype(np.float32), axis=1)
    d[:, 0] = 0.0
    sig, rho = oracle.smooth_knn_dist(d, 15.0, 1.0, include_self=True, fp64_sums=True)
    act = d[:, 1:].astype(np.float64)
    psum = np.where(act - rho[:, None] > 0, np.exp(-(act - rho[:, None]) / sig[:, None]), 1.0).sum(1)
    assert np.all(np.abs(psum - np.log2(15.0)) < 1.2e-5)
    s32, r32 = oracle.smooth_knn_dist(d, 15.0, 1.0, include_self=True)
    assert np.array_equal(rho, r32) and np.max(np.abs(sig - s32) / s32) < 2e-4
    with_self = np.array([[0.0, 0.2, 0.4, 0.8], [0.0, 0.1, 0.5, 0.9]], np.float32)
    s1, r1 = oracle.smooth_knn_dist(with_self, 3.0, 1.0, include_self=True, fp64_sums=True)
    s2, r2 = oracle.smooth_knn_dist(with_self[:, 1:], 3.0, 1.0, include_self=False,
                                    fp64_sums=True)
    assert np.array_equal(s1, s2) and np.array_equal(r1, r2)
    # two neighbours: psum = 1 + exp(-(d2 - d1) / sigma) = log2(3) in closed form
    row = np.array([[0.0, 0.3, 1.1]], np.float32)
    s, r = oracle.smooth_knn_dist(row, 3.0, 1.0, include_self=True, fp64_sums=True)
    closed = (1.1 - 0.3) / -np.log(np.log2(3.0) - 1.0)
    assert r[0] == np.float32(0.3) and abs(s[0] - closed) / closed < 1e-4


def test_distance_closed_forms(oracle):
    x = np.array([1.0, -2.0, 0.0, 4.0], np.float32)
    y = np.array([-1.0, 2.0, 0.0, 1.0], np.float32)
    assert oracle.sq_euclidean(x, y) == np.float32(4 + 16 + 0 + 9)
    c = oracle.cosine_distance(x, y)
    expected = 1.0 - (x.astype(np.float64) @ y) / (np.linalg.norm(x) * np.linalg.norm(y))
    assert abs(c - expected) < 1e-6
    z = np.zeros(4, np.float32)
    assert oracle.cosine_distance(z, z) == 0.0 and oracle.cosine_distance(z, y) == 1.0


def test_eight_lane_order_differs_from_sequential_only_by_rounding(oracle, rng):
    # d >= 64 uses 8 partial sums (lib.rs:2470-2492); check against a numpy
    # restatement of that exact order in float32
    x = rng.normal(size=100).astype(np.float32)
    y = rng.normal(size=100).astype(np.float32)
    lanes = np.zeros(8, np.float32)
    for i in range(0, 96, 8):
        df = (x[i:i + 8] - y[i:i + 8]).astype(np.float32)
        lanes = (lanes + df * df).astype(np.float32)
    s = np.float32(0.0)
    for v in lanes:
        s = np.float32(s + v)
    for i in range(96, 100):
        df = np.float32(x[i] - y[i])
        s = np.float32(s + df * df)
    assert oracle.sq_euclidean(x, y) == s


def test_tie_break_kat(oracle):
    # sparse.rs:1070-1122: four equidistant points, order must be by index
    ref = np.array([[1.0, 0.0], [-1.0, 0.0], [0.0, 1.0], [0.0, -1.0]], np.float32)
    q = np.zeros((1, 2), np.float32)
    for metric in ("euclidean",):
        idx, dist = oracle.knn_to_reference(q, ref, 3, metric)
        assert idx[0].tolist() == [0, 1, 2]
        assert np.all(dist == 1.0)


@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
@pytest.mark.parametrize("n,d", [(400, 16), (500, 70), (300, 128)])
def test_oracle_knn_matches_fp64(oracle, rng, metric, n, d):
    x = rng.normal(size=(n, d)).astype(np.float32)
    idx, dist = oracle.knn(x, 15, metric)
    i64, d64 = knn_fp64(x, None, 15, metric)
    # self first, sorted
    assert np.all(np.diff(dist, axis=1) >= 0)
    if metric == "euclidean":
        assert np.all(idx[:, 0] == np.arange(n))
    diff_rows = np.where(np.any(idx != i64, axis=1))[0]
    for r in diff_rows:
        # only near-ties may swap: compare the FP64 distances of both sets
        assert abs(np.sort(d64[r])[-1] - dist[r, -1]) <= 1e-5 * max(1.0, dist[r, -1])
    assert len(diff_rows) <= max(1, n // 100)
    # relative 1e-5, absolute 1e-6 near zero (cosine self-distance is ~1e-7 in f32)
    assert np.all(np.abs(dist - d64) <= 1e-5 * d64 + 1e-6)


def test_membership_and_union_closed_forms(oracle):
    # 3 points, k = 3 (self + 2): hand-made kNN rows
    idx = np.array([[0, 1, 2], [1, 0, 2], [2, 1, 0]], np.int64)
    dist = np.array([[0.0, 1.0, 2.0], [0.0, 1.0, 1.5], [0.0, 1.5, 2.0]], np.float32)
    sig = np.array([0.5, 0.25, 1.0], np.float32)
    rho = np.array([1.0, 1.0, 1.5], np.float32)
    h, t, w = oracle.fuzzy_graph(idx, dist, sig, rho, 1.0, 0)
    wd = {}
    for i in range(3):
        for c in range(1, 3):
            j = idx[i, c]
            v = 1.0 if dist[i, c] - rho[i] <= 0 else np.exp(-(dist[i, c] - rho[i]) / sig[i])
            wd[(i, j)] = np.float32(v)
    got = {(int(a), int(b)): float(c) for a, b, c in zip(h, t, w)}
    for (i, j), v in got.items():
        a = wd.get((i, j), 0.0)
        b = wd.get((j, i), 0.0)
        assert abs(v - (a + b - a * b)) <= 1e-6
        assert got[(j, i)] == v
    assert list(h) == sorted(h)


def test_schedule_counts_closed_form(oracle):
    # an edge of weight w fires floor((T - 1) * w / w_max / 1)... check directly:
    # eps = T / (T * w / wmax); with w = wmax it fires every epoch from 1
    w = np.array([1.0, 0.5, 0.25, 0.1], np.float32)
    pos, neg, act, negs = oracle.schedule_counts(w, 200, 5)
    assert pos[0] == 199 and pos[1] == 99 and pos[2] == 49 and pos[3] == 19
    # negatives: floor((t - eonns) / (eps/5)) accumulated; total = 5 * pos
    # minus the partial first interval
    assert np.all(neg <= 5 * pos) and np.all(neg >= 5 * pos - 5)
    assert act.sum() == pos.sum() and negs.sum() == neg.sum()
    assert act[0] == 0  # nothing fires at epoch 0


def test_density_start_epoch(oracle):
    assert oracle.density_start_epoch(500, 0.3) == 350
    assert oracle.density_start_epoch(10, 0.0) == 9
    assert oracle.density_start_epoch(200, 1.0) == 0


def test_spectral_init_matches_eigh_on_connected_graph(oracle):
    # the iterative path (d > 8) must land on eigenvectors 2..3 of
    # I + D^-1/2 W D^-1/2; compare subspaces (signs / rotation are free)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(400, 12)).astype(np.float32)
    idx, dist = oracle.knn(x, 15)
    sig, rho = oracle.smooth_knn_dist(dist, 15.0)
    h, t, w = oracle.fuzzy_graph(idx, dist, sig, rho, 1.0, 500)
    nc, _ = oracle.connected_components(400, h, t)
    assert nc == 1
    y, ok = oracle.initialize_embedding(x, 2, h, t, w, "spectral", 42)
    assert ok
    W = np.zeros((400, 400))
    W[h, t] = w
    W = np.maximum(W, W.T)
    deg = W.sum(1)
    dinv = 1 / np.sqrt(deg)
    M = np.eye(400) + dinv[:, None] * W * dinv[None, :]
    ev, vec = np.linalg.eigh(M)
    top = vec[:, np.argsort(ev)[::-1][1:3]]
    q, _ = np.linalg.qr(y.astype(np.float64) - 0.0)
    s = np.linalg.svd(top.T @ q, compute_uv=False)
    assert s.min() > 0.99


def test_components_ordered_by_smallest_member(oracle):
    h = np.array([0, 3, 1, 4, 2, 5], np.int64)
    t = np.array([3, 0, 4, 1, 5, 2], np.int64)
    nc, lab = oracle.connected_components(7, h, t)
    assert nc == 4
    assert lab.tolist() == [0, 1, 2, 0, 1, 2, 3]


def test_edge_laplacian_equals_dense_affinity_laplacian(oracle, rng):
    # lib.rs:7457-7538: the edge-list Laplacian equals the one built from the
    # dense, max-symmetrised affinity matrix (to 1e-12), on the reference's
    # 5-node graph and on a random one; and both equal I - D^-1/2 W D^-1/2
    # formed here in numpy
    graphs = [(5, [0, 1, 1, 2, 3, 4, 0], [1, 0, 2, 3, 4, 3, 4],
               [0.4, 0.7, 0.8, 0.6, 0.2, 0.5, 0.3])]
    n = 40
    h = rng.integers(0, n, 200)
    t = rng.integers(0, n, 200)
    graphs.append((n, h.tolist(), t.tolist(), rng.uniform(0, 1, 200).astype(np.float32).tolist()))
    for n, h, t, w in graphs:
        a = oracle.normalized_laplacian(n, h, t, w)
        b = oracle.normalized_laplacian(n, h, t, w, from_affinity=True)
        assert np.abs(a - b).max() < 1e-12
        W = np.zeros((n, n))
        for hh, tt, ww in zip(h, t, w):
            if hh != tt:
                W[hh, tt] = max(W[hh, tt], float(np.float32(ww)))
        W = np.maximum(W, W.T)
        deg = W.sum(1)
        inv = np.where(deg > 0, 1.0 / np.sqrt(np.where(deg > 0, deg, 1.0)), 0.0)
        L = np.eye(n) - inv[:, None] * W * inv[None, :]
        assert np.abs(a - L).max() < 1e-12


def test_densmap_zero_lambda_is_the_standard_embedding(oracle):
    # lib.rs:6078-6106: densMAP with lambda = 0 leaves the embedding unchanged
    x = np.random.default_rng(3).normal(size=(400, 12)).astype(np.float32)
    plain = oracle.fit(x, oracle.OracleParams(n_epochs=60, random_seed=9)).embedding
    dm0 = oracle.fit(x, oracle.OracleParams(n_epochs=60, random_seed=9, densmap=True,
                                            dens_lambda=0.0)).embedding
    assert np.array_equal(plain, dm0)
    dm = oracle.fit(x, oracle.OracleParams(n_epochs=60, random_seed=9, densmap=True,
                                           dens_lambda=2.0)).embedding
    assert not np.array_equal(plain, dm)


# The generators behind rand 0.8.5's SmallRng (xoshiro256++, seeded through
# SplitMix64), against the published reference outputs: xoshiro256++ from the
# state {1, 2, 3, 4} (Blackman & Vigna's xoshiro256plusplus.c, the vector
# rand_xoshiro tests against; the first two outputs also follow by hand:
# rotl(1 + 4, 23) + 1 = 41943041, then 58720359), and SplitMix64 from seed
# 1234567 (Vigna's splitmix64.c).
XOSHIRO256PP_1234 = [41943041, 58720359, 3588806011781223, 3591011842654386,
                     9228616714210784205, 9973669472204895162, 14011001112246962877,
                     12406186145184390807, 15849039046786891736, 10450023813501588000]
SPLITMIX64_1234567 = [6457827717110365317, 3203168211198807973, 9817491932198370423,
                      4593380528125082431, 16408922859458223821]


def test_xoshiro256pp_reference_vector(oracle):
    assert [int(v) for v in oracle.xoshiro256pp([1, 2, 3, 4], 10)] == XOSHIRO256PP_1234


def test_splitmix64_reference_vector(oracle):
    assert [int(v) for v in oracle.splitmix64(1234567, 5)] == SPLITMIX64_1234567


def test_smallrng_seeding_composes_splitmix_and_xoshiro(oracle):
    # seed_from_u64: the state is the first four SplitMix64 outputs
    for seed in (0, 1, 42, 0x7A5B3D91CE42A117):
        st = oracle.splitmix64(seed, 4)
        assert np.array_equal(oracle.smallrng_u64(seed, 16), oracle.xoshiro256pp(st, 16))


def test_gen_index_is_lemire_widening_multiply(oracle):
    # rand 0.8.5 UniformInt::sample_single for u64: v * range as a 128-bit
    # product, accept when its low half is at most (range << lz(range)) - 1,
    # return the high half; checked on the raw stream in Python integers
    for seed, rng_ in ((3, 10), (7, 1000003), (11, 1 << 20), (13, 3 * (1 << 40) + 1)):
        raw = [int(v) for v in oracle.smallrng_u64(seed, 400)]
        zone = ((rng_ << (64 - rng_.bit_length())) - 1) & ((1 << 64) - 1)
        want, i = [], 0
        while len(want) < 100:
            m = raw[i] * rng_
            i += 1
            if (m & ((1 << 64) - 1)) <= zone:
                want.append(m >> 64)
        assert [int(v) for v in oracle.smallrng_index(seed, rng_, 100)] == want
