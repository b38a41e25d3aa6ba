"""The oracle against the reference's own published whole-fit quality figures
(its CI smoke gates on digits and breast_cancer: benchmarks/ann-e2e-smoke.eval.json,
benchmarks/consistency-smoke.eval.json), under the reference's protocol
(tests/smoke_protocol.py).  These figures are the only reference outputs that
reach the stages fed by its random streams (random init, SGD negatives), so
they pin the oracle's statistical behaviour where no bit-level vector can.

The reference publishes one seed-42 run per figure; the oracle's mean over
ten seeds must lie within 0.01 of it.  The fixture's per-seed figures must be
what the oracle computes now."""
import json
import os

import numpy as np
import pytest

from tests import smoke_protocol as P

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "reference_smoke.json")))


@pytest.mark.parametrize("gate", sorted(P.GATES))
@pytest.mark.parametrize("name", ["digits", "breast_cancer"])
def test_oracle_within_001_of_reference_smoke_figures(oracle, gate, name):
    cap, pre = P.GATES[gate]
    rec = FIX["gates"][gate]["datasets"][name]
    ref = P.REFERENCE[gate][name]
    assert rec["reference"] == ref
    x = P.load_dataset(name)
    idx, dist = P.shared_exact_knn(x)
    ts, rs = [], []
    for s in FIX["seeds"]:
        prm = oracle.OracleParams(random_seed=s, n_epochs=FIX["n_epochs"], init="random")
        y = (oracle.fit(x, prm, knn_idx=idx, knn_dist=dist) if pre else oracle.fit(x, prm)).embedding
        t, r = P.consistency(x, y, idx, cap)
        ts.append(t)
        rs.append(r)
    np.testing.assert_allclose(ts, rec["oracle_trust"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(rs, rec["oracle_recall"], rtol=0, atol=1e-9)
    assert abs(np.mean(ts) - ref["trust"]) <= 0.01, (np.mean(ts), ref)
    assert abs(np.mean(rs) - ref["recall"]) <= 0.01, (np.mean(rs), ref)


def test_smoke_protocol_keeps_the_reference_quirks():
    """The protocol's two easily-lost details: the shared kNN drops self, and
    the embedding-side kNN drops the nearest non-self neighbour (ranks
    2..k+1), as compare_real_impls_fair.py:457-461 does."""
    x = np.array([[0.0], [1.0], [3.0], [7.0], [15.0]], np.float32)
    idx, dist = P.shared_exact_knn(x, k=2)
    assert idx.tolist() == [[1, 2], [0, 2], [1, 0], [2, 1], [3, 2]]
    assert dist[0].tolist() == [1.0, 3.0]
    e = P._knn_indices(x, 2)
    assert e.shape == (5, 2)
    # ranks 2..3 of the embedding: row 0's non-self order is 1, 2, 3
    assert e[0].tolist() == [2, 3]
