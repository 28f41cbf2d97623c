"""Propagation rounds replayed from a captured hipGraph (option graph_rounds).

The first calculate() on a data set probes the round mode; later calls replay
all rounds from one graph, with the forest and leaf stages running eagerly in
between.  Every replay must store the graph the round-by-round launches store,
bit for bit, in every mode: exact rounds, screened rounds, COSINE parity.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.mark.parametrize("metric,opts", [(0, {}), (0, {"sketch": 1}), (1, {"sketch": 1}), (2, {}),
                                         (2, {"cosine_fixed": 1, "sketch": 1}), (0, {"split": 1, "sketch": 1})])
def test_graph_replays_match_eager_rounds(aknn, datasets, metric, opts):
    x = datasets.gmm_c1(6000, 24, clusters=6, seed=5)[0]
    print(f"[gpu] graph rounds metric={metric} {opts}", flush=True)
    ref = aknn.Graph()
    ref.initialize(x, 8, 3, 24, 3, metric, graph_rounds=0, **opts)
    ref.calculate()
    want = ref.graph()
    ref.close()
    g = aknn.Graph()
    g.initialize(x, 8, 3, 24, 3, metric, graph_rounds=1, **opts)
    replays = 0
    for step in range(4):
        g.calculate()
        ids, vals = g.graph()
        assert (ids == want[0]).all() and (bits(vals) == bits(want[1])).all(), f"call {step}"
        replays += g.kernel_stats("propagate_graph")["launches"] > 0
    # the first call may probe the mode; the later ones replay the graph
    assert replays >= 3
    g.close()
