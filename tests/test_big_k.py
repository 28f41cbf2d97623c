"""Lists longer than 64 (64 < k <= 256; big_k.hip).

The leaf stage keys each tree's start ranges into a table and merges every
point's row with its table row by sorting (row + a chunk of candidates) and
keeping the first k distinct pairs; the propagation rounds do the same with
the two-hop candidates.  The graphs must be the oracle's bit for bit, which
the oracle computes for any k (euclideandistancematrixgpu.cu:2768 sizes the
reference's scratch for any k).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.mark.parametrize("metric,k,opts", [
    (0, 65, {}), (0, 100, {}), (0, 128, {"leaf_mfma": 1, "sketch": 1}), (0, 256, {}),
    (1, 100, {}), (3, 100, {}), (2, 100, {"cosine_fixed": 1}),
])
def test_big_k_bit_exact(aknn, oracle, datasets, metric, k, opts):
    x = datasets.gmm_c1(2500, 20, clusters=6, seed=13)[0]
    trees, leaf, props = 3, 2 * k, 2
    m = metric if metric < 3 else 1
    extra = {"taxicab_l1": 1} if metric == 3 else {}
    print(f"[gpu] big k: metric={metric} k={k} {opts}", flush=True)
    g = aknn.Graph()
    g.initialize(x, k, trees, leaf, props, m, **opts, **extra)
    g.calculate()
    ids, vals = g.graph()
    g.close()
    om = oracle.TAXICAB_L1 if metric == 3 else m
    _, oids, ovals, _ = oracle.run_ex(x, k, trees, leaf, props, om, cosine_fixed=opts.get("cosine_fixed", 0))
    assert (ids == oids).all() and (bits(vals) == bits(ovals)).all()
    # every row: distinct ids other than the point, sorted by (value, id)
    rows = np.arange(len(x))[:, None]
    assert (ids != rows).all()
    s = np.sort(ids, axis=1)
    assert ((s[:, 1:] != s[:, :-1]) | (s[:, 1:] < 0)).all()


def test_big_k_median_forest_and_small_leaves(aknn, oracle, datasets):
    """Median-split trees, leaves smaller than k + 1 (start ranges from
    ancestors), stubs left in rows when a point's ranges hold fewer than k
    others."""
    x = datasets.gmm_c1(1500, 12, clusters=4, seed=3)[0]
    k = 90
    print("[gpu] big k: median forest, small leaves", flush=True)
    for split, leaf, props in ((1, 40, 1), (0, 16, 0), (0, 60, 3)):
        g = aknn.Graph()
        g.initialize(x, k, 2, leaf, props, 0, split=split)
        g.calculate()
        ids, vals = g.graph()
        g.close()
        _, oids, ovals, _ = oracle.run_ex(x, k, 2, leaf, props, 0, split=split)
        assert (ids == oids).all() and (bits(vals) == bits(ovals)).all(), (split, leaf, props)


def test_big_k_cosine_parity_refused(aknn, datasets):
    x = datasets.gmm_c1(500, 8, clusters=3, seed=1)[0]
    g = aknn.Graph()
    with pytest.raises(RuntimeError, match="cosine_fixed"):
        g.initialize(x, 80, 2, 100, 1, 2)
    g.close()
