"""aknn_recall_noalloc (no device allocation, static scratch) against
aknn_recall through the C ABI: the same recall, bit for bit, on consecutive
calls that reuse the scratch with smaller inputs, and the size limit refused
with an error before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graph(aknn, x, k, trees, leaf, props, metric):
    print(f"[no_alloc] graph n={len(x)} d={x.shape[1]} k={k} metric={metric}", flush=True)
    g = aknn.Graph()
    g.initialize(x, k, trees, leaf, props, metric)
    g.calculate()
    return g


def test_recall_noalloc_matches_recall(aknn, datasets):
    rng = np.random.default_rng(77)
    xa = datasets.gmm_c1(8000, 24, clusters=6, seed=21)[0]
    ga = _graph(aknn, xa, 10, 3, 32, 1, aknn.EUCLIDEAN)
    try:
        for nq in (2500, 400):
            q = np.sort(rng.choice(len(xa), size=nq, replace=False))
            print(f"[no_alloc] EUCLIDEAN nq={nq}", flush=True)
            want = ga.recall(q)
            got = ga.recall_noalloc(q)
            assert 0.0 < want <= 1.0
            assert got == want, (got, want)
        # a query count whose buffers exceed the static scratch is refused
        with pytest.raises(aknn.AknnError, match="static scratch"):
            ga.recall_noalloc(np.zeros(400_000, np.int64))
        with pytest.raises(aknn.AknnError, match="out of range"):
            ga.recall_noalloc(np.array([0, len(xa)], np.int64))
    finally:
        ga.close()
    xb = datasets.gmm_c1(3000, 40, clusters=4, seed=22)[0]
    gb = _graph(aknn, xb, 12, 2, 24, 1, aknn.COSINE)
    try:
        q = np.sort(rng.choice(len(xb), size=300, replace=False))
        print("[no_alloc] COSINE nq=300", flush=True)
        want = gb.recall(q)
        got = gb.recall_noalloc(q)
        assert 0.0 < want <= 1.0
        assert got == want, (got, want)
    finally:
        gb.close()
