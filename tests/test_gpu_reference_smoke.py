"""The GPU fit under the reference's own CI smoke protocol (digits and
breast_cancer, k = 15, 2-D, 200 epochs, random init; tests/smoke_protocol.py):
its mean trustworthiness@15 and original-kNN recall@15 over the oracle's ten
seeds lie within 0.01 of the oracle's means, both ways.  The oracle's means
are themselves held within 0.01 of the reference's published figures
(tests/test_reference_smoke.py), so this ties the GPU's random-stream stages
(random init, Hogwild SGD with hashed negatives) to the reference's outputs."""
import json
import os

import numpy as np
import pytest

from tests import smoke_protocol as P
from umapers_amd import Umap

pytestmark = pytest.mark.gpu

FIX = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                  "reference_smoke.json")))


@pytest.mark.parametrize("gate", sorted(P.GATES))
@pytest.mark.parametrize("name", ["digits", "breast_cancer"])
def test_gpu_within_001_of_oracle_smoke_figures(gpu, gate, name):
    cap, pre = P.GATES[gate]
    rec = FIX["gates"][gate]["datasets"][name]
    x = P.load_dataset(name)
    idx, dist = P.shared_exact_knn(x)
    ts, rs = [], []
    for s in FIX["seeds"]:
        m = Umap(n_neighbors=15, n_components=2, n_epochs=FIX["n_epochs"], init="random",
                 random_seed=s)
        y = m.fit_transform_with_knn(x, idx, dist) if pre else m.fit_transform(x)
        t, r = P.consistency(x, y, idx, cap)
        ts.append(t)
        rs.append(r)
    print(f"{gate} {name}: gpu trust {np.mean(ts):.4f} recall {np.mean(rs):.4f}; oracle "
          f"{rec['oracle_trust_mean']:.4f} / {rec['oracle_recall_mean']:.4f}; reference "
          f"{rec['reference']['trust']:.4f} / {rec['reference']['recall']:.4f}", flush=True)
    assert abs(np.mean(ts) - rec["oracle_trust_mean"]) <= 0.01, (ts, rec["oracle_trust"])
    assert abs(np.mean(rs) - rec["oracle_recall_mean"]) <= 0.01, (rs, rec["oracle_recall"])
