"""Fixed-mode semantics: the reference's quirks corrected, next to parity.

  taxicab_l1  TAXICAB keys and stores the L1 total; the reference stores its
              sqrtf (euclideandistancematrixgpu.cu:568, :1477);
  scale_all   normalization divides every dimension by the max-abs over all
              dimensions; the reference covers only the first 256
              (.cu:57-65, :76-84);
  cosine_fixed (already covered by the parity tests) carries COSINE
              distances instead of the 1 - key flip.

The oracle's fixed modes are pinned against closed forms and independent
restatements; the GPU (marked gpu) against the oracle, bit for bit.
"""
import numpy as np
import pytest


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def test_l1_distance_is_the_chain_the_reference_roots(oracle):
    a = np.array([0, 0], np.float32)
    b = np.array([3, 4], np.float32)
    assert oracle.distance(a, b, oracle.TAXICAB_L1) == np.float32(7.0)
    rng = np.random.default_rng(3)
    for d in (1, 7, 256, 300, 600):
        x = rng.normal(size=(2, d)).astype(np.float32)
        l1 = oracle.distance(x[0], x[1], oracle.TAXICAB_L1)
        # the reference's value is sqrtf of the same FP32 total
        assert oracle.distance(x[0], x[1], oracle.TAXICAB) == np.sqrt(l1, dtype=np.float32)
        # the chunked FP32 sum (256-dimension chunks, each from 0)
        tot = np.float32(0)
        for c0 in range(0, d, 256):
            p = np.float32(0)
            for v in np.abs(x[0, c0:c0 + 256] - x[1, c0:c0 + 256]):
                p = np.float32(p + v)
            tot = np.float32(tot + p)
        assert l1 == tot


@pytest.mark.parametrize("shape", [(50, 300), (40, 700), (30, 8)])
def test_scale_all_vs_numpy(oracle, shape):
    rng = np.random.default_rng(6)
    x = (rng.normal(3, 2, shape) * 10).astype(np.float32)
    got = oracle.normalize(x, scale_all=1)
    n, d = shape
    want = x.copy()
    for j in range(d):
        m = np.cumsum(want[:, j], dtype=np.float32)[-1] / np.float32(n)
        want[:, j] = want[:, j] - np.float32(m)
    want = want / np.abs(want).max()
    assert (bits(got) == bits(want)).all()
    assert np.abs(got).max() == 1.0
    if d <= 256:
        assert (bits(got) == bits(oracle.normalize(x))).all()


@pytest.mark.parametrize("n,d", [(5000, 3), (4097, 5), (2048, 2), (700, 300)])
def test_mean_fp64_vs_python(oracle, n, d):
    """mean_fp64: FP64 sums over blocks of 2048 points in order, block sums in
    order, rounded once (restated here in pure Python floats, which are FP64),
    then the reference's centring and max-abs scaling."""
    rng = np.random.default_rng(n + d)
    x = (rng.normal(2, 3, (n, d)) * 10).astype(np.float32)
    got = oracle.normalize(x, mean_fp64=1)
    want = x.copy()
    for j in range(d):
        tot = 0.0
        for b in range(0, n, 2048):
            part = 0.0
            for v in x[b:b + 2048, j].tolist():
                part = part + v
            tot = tot + part
        want[:, j] = want[:, j] - np.float32(tot / n)
    dm = min(d, 256)
    want[:, :dm] = want[:, :dm] / np.abs(want[:, :dm]).max()
    assert (bits(got) == bits(want)).all()


def test_mean_fp64_integer_data_is_exact(oracle):
    """Integer values: every FP64 partial sum is exact, so the mean is the
    exact mean rounded once, whatever the blocking (a dropped block or a
    wrong divisor changes it)."""
    rng = np.random.default_rng(2)
    n, d = 6000, 4
    x = rng.integers(0, 200000, (n, d)).astype(np.float32)  # sums pass 2^24: the FP32 chain rounds
    got = oracle.normalize(x, mean_fp64=1)
    exact = np.float32(x.astype(np.int64).sum(0) / n)
    want = x - exact
    want = want / np.abs(want).max()
    assert (bits(got) == bits(want)).all()
    assert not (bits(got) == bits(oracle.normalize(x))).all()  # the sequential FP32 mean differs here


def test_l1_pipeline_brute_force(oracle, datasets):
    """One tree with a leaf bound >= N: every point's candidates are all the
    others, so the initial graph is the brute-force top-k by (L1, id)."""
    x = datasets.gmm_c1(400, 6, clusters=3, seed=9)[0]
    k = 7
    xn, ids, vals, _ = oracle.run_ex(x, k, 1, 1000, 0, oracle.TAXICAB_L1)
    for i in range(0, 400, 13):
        keys = [(oracle.distance(xn[i], xn[j], oracle.TAXICAB_L1), j) for j in range(400) if j != i]
        keys.sort()
        assert ids[i].tolist() == [j for _, j in keys[:k]]
        assert (bits(vals[i]) == bits([v for v, _ in keys[:k]])).all()


def test_l1_and_root_rank_alike(oracle, datasets):
    """sqrt is monotone: the L1 graph's neighbours are the reference's up to
    ties that sqrtf creates; the stored values are the squares' roots."""
    x = datasets.gmm_c1(1500, 10, clusters=4, seed=2)[0]
    _, i1, v1, _ = oracle.run_ex(x, 8, 3, 24, 1, oracle.TAXICAB_L1)
    _, i2, v2, _ = oracle.run_ex(x, 8, 3, 24, 1, oracle.TAXICAB)
    same = (i1 == i2).all(1).mean()
    assert same > 0.98
    rows = (i1 == i2).all(1)
    assert (bits(np.sqrt(v1[rows], dtype=np.float32)) == bits(v2[rows])).all()


# ---- the HIP build against the oracle ----------------------------------------

def _gpu(aknn, x, k, trees, leaf, props, metric, **opts):
    print(f"[gpu] fixed n={len(x)} d={x.shape[1]} k={k} T={trees} L={leaf} P={props} metric={metric} {opts}",
          flush=True)
    g = aknn.Graph()
    g.initialize(x, k, trees, leaf, props, metric, **opts)
    g.calculate()
    return g


@pytest.mark.gpu
@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("sketch", [0, 1])
def test_gpu_taxicab_l1_bit_exact(aknn, oracle, datasets, split, sketch):
    x = datasets.gmm_c1(3000, 24, clusters=6, seed=3)[0]
    g = _gpu(aknn, x, 8, 4, 20, 2, 1, taxicab_l1=1, split=split, sketch=sketch)
    xn, ids, vals, _ = oracle.run_ex(x, 8, 4, 20, 2, oracle.TAXICAB_L1, split=split)
    gi, gv = g.graph()
    assert (gi == ids).all() and (bits(gv) == bits(vals)).all()


@pytest.mark.gpu
@pytest.mark.parametrize("metric", [0, 1, 2])
def test_gpu_fixed_mode_wide_rows(aknn, oracle, datasets, metric):
    """D=300: scale_all changes the dimensions past 256; all fixed semantics at once."""
    x = datasets.sparse_c3(1200, 300, clusters=8, seed=6)
    g = _gpu(aknn, x, 5, 2, 40, 1, metric, fixed=1)
    om = oracle.TAXICAB_L1 if metric == 1 else metric
    xn, ids, vals, _ = oracle.run_ex(x, 5, 2, 40, 1, om, cosine_fixed=1, scale_all=1)
    assert (bits(g.normalized()) == bits(xn)).all()
    assert np.abs(xn).max() == 1.0
    gi, gv = g.graph()
    assert (gi == ids).all() and (bits(gv) == bits(vals)).all()


@pytest.mark.gpu
@pytest.mark.parametrize("n,d", [(5000, 24), (70000, 128), (3000, 300)])
def test_gpu_mean_fp64_bit_exact(aknn, oracle, datasets, n, d):
    """The GPU's blocked FP64 mean and normalization equal the oracle's."""
    x = datasets.gmm_c1(n, d, clusters=5, seed=90 + d)[0]
    g = aknn.Graph()
    g.initialize(x, 6, 2, 40, 0, 0, mean_fp64=1)
    g.calculate()
    assert (bits(g.normalized()) == bits(oracle.normalize(x, mean_fp64=1))).all()
    g.close()
