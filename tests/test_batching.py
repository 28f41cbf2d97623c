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


@pytest.mark.parametrize("metric,opts", [(0, {}), (0, {"leaf_mfma": 1, "sketch": 1}), (1, {}),
                                         (2, {"cosine_fixed": 1, "leaf_mfma": 1}), (3, {}),
                                         (0, {"leaf_mfma": 1, "split": 1}), (0, {"leaf_mfma": 1, "k": 40})])
def test_all_trees_launch_matches_per_tree(aknn, datasets, oracle, metric, opts):
    """Every tree's leaf k-NN in one persistent launch (option leaf_batch)
    stores the graph of one launch per tree, which is the oracle's."""
    opts = dict(opts)
    k = opts.pop("k", 10)
    m = metric if metric < 3 else 1
    extra = {"taxicab_l1": 1} if metric == 3 else {}
    x = datasets.gmm_c1(5000, 20, clusters=8, seed=9)[0]
    print(f"[gpu] all-trees leaf launch metric={metric} k={k} {opts}", flush=True)
    graphs, stats = [], []
    for batch in (0, 1):
        g = aknn.Graph()
        g.initialize(x, k, 5, 24, 0, m, leaf_batch=batch, **opts, **extra)
        g.calculate()
        graphs.append(g.graph())
        stats.append((g.kernel_stats("leaf_knn"), g.kernel_stats("leaf_mfma"), g.kernel_stats("leaf_exact")))
        g.close()
    assert (graphs[0][0] == graphs[1][0]).all() and (bits(graphs[0][1]) == bits(graphs[1][1])).all()
    assert stats[0][0]["launches"] == 5 and stats[1][0]["launches"] == 1
    # the same trees counted as screened and exact either way, each with a time
    assert stats[1][1]["launches"] + stats[1][2]["launches"] == 5
    if opts.get("leaf_mfma") and metric in (0, 2):
        assert stats[1][1]["launches"] >= 1 and stats[1][1]["total_ms"] > 0
    if metric in (0, 1) and not opts.get("split"):
        _, oids, ovals, _ = oracle.run(x, k, 5, 24, 0, m)
        assert (graphs[1][0] == oids).all() and (bits(graphs[1][1]) == bits(ovals)).all()


def test_all_trees_launch_switches_screen_off(aknn, datasets):
    """Uniform data in many dimensions: the screen keeps too many pairs after
    the first screened tree, and the launch's last workgroup at the barrier
    switches the remaining trees to the exact body (same graph)."""
    rng = np.random.default_rng(3)
    x = rng.random((4000, 300), dtype=np.float32)
    print("[gpu] all-trees leaf launch: screen switch", flush=True)
    out = []
    for batch in (0, 1):
        g = aknn.Graph()
        g.initialize(x, 16, 6, 64, 0, 0, leaf_batch=batch, leaf_mfma=1)
        g.calculate()
        out.append((g.graph(), g.kernel_stats("leaf_mfma")["launches"]))
        g.close()
    assert (out[0][0][0] == out[1][0][0]).all() and (bits(out[0][0][1]) == bits(out[1][0][1])).all()
    # the device decides from the tree just finished, the host from one two
    # trees back: the device switches no later
    assert 1 <= out[1][1] <= out[0][1]


@pytest.mark.parametrize("opts", [{}, {"leaf_mfma": 1}, {"leaf_mfma": 1, "split": 1}])
def test_leaf_launch_form_tuned_per_data_set(aknn, datasets, opts):
    """leaf_batch 2 (the default): the first calculate() runs both launch
    forms (the all-trees launch, then the rows again from stubs one launch per
    tree) and keeps the faster; every call stores the graph of one launch per
    tree."""
    x = datasets.gmm_c1(6000, 24, clusters=6, seed=2)[0]
    print(f"[gpu] leaf launch form tuning {opts}", flush=True)
    ref = aknn.Graph()
    ref.initialize(x, 10, 4, 32, 1, 0, leaf_batch=0, **opts)
    ref.calculate()
    want = ref.graph()
    ref.close()
    g = aknn.Graph()
    g.initialize(x, 10, 4, 32, 1, 0, **opts)
    forms = []
    for step in range(3):
        g.calculate()
        ids, vals = g.graph()
        assert (ids == want[0]).all() and (bits(vals) == bits(want[1])).all(), f"call {step}"
        f = g.leaf_form()
        forms.append(f["form"])
        assert f["tuning_ms"]["per_tree"] > 0 and f["tuning_ms"]["all_trees"] > 0
    # the learning call ends with the launches per tree; the later ones keep the choice
    assert forms[0] == "per_tree" and forms[1] == forms[2]
    want_form = "all_trees" if f["tuning_ms"]["all_trees"] < f["tuning_ms"]["per_tree"] else "per_tree"
    assert forms[2] == want_form
    # an option change learns again
    g.set_option("leaf_batch", 2)
    g.calculate()
    assert g.leaf_form()["form"] == "per_tree"
    ids, vals = g.graph()
    assert (ids == want[0]).all() and (bits(vals) == bits(want[1])).all()
    g.close()
