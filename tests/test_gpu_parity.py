"""GPU path vs the CPU oracle, bit for bit, through the C ABI.

Every case compares the HIP build with the oracle on identical seeded inputs
at sizes the oracle finishes in seconds: normalized points, every tree's leaf
order and start ranges (and the rand() iterations it consumed), and the graph
ids and FP32 values.  Integer outputs and float values are compared exactly.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR, TESTS_DIR

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def gpu_run(aknn, x, k, trees, leaf, props, metric, **opts):
    print(f"[gpu] n={len(x)} d={x.shape[1]} k={k} T={trees} L={leaf} P={props} metric={metric} {opts}", flush=True)
    g = aknn.Graph()
    g.initialize(x, k, trees, leaf, props, metric, **opts)
    g.calculate()
    return g


def check_forest(oracle, g, x_norm, trees, leaf, k, seed=1, order=0):
    perm, rs, rc, info = oracle.build_forest(x_norm, trees, leaf, k, seed=seed, rng_order=order)
    for t in range(trees):
        gp, grs, grc, it = g.forest(t)
        assert (gp == perm[t]).all(), f"tree {t} leaf order"
        assert (grs == rs[t]).all() and (grc == rc[t]).all(), f"tree {t} start ranges"
        assert it == info[t][0], f"tree {t} iterations"


def check_graph(oracle, g, x_raw, k, trees, leaf, props, metric, seed=1, order=0, cosine_fixed=0):
    xn, ids, vals, _ = oracle.run(x_raw, k, trees, leaf, props, metric, seed=seed, rng_order=order,
                                  cosine_fixed=cosine_fixed)
    gx = g.normalized()
    assert (bits(gx) == bits(xn)).all(), "normalized points"
    gi, gv = g.graph()
    bad = np.nonzero((gi != ids).any(1) | (bits(gv) != bits(vals)).any(1))[0]
    assert len(bad) == 0, f"{len(bad)} rows differ, first {bad[:5]}: gpu {gi[bad[0]]} {gv[bad[0]]} oracle {ids[bad[0]]} {vals[bad[0]]}"
    return xn


def test_readme_example_cli(tmp_path):
    """The reference's README example through the drop-in CLI: output bytes match the exact-arithmetic KAT."""
    kat = json.load(open(os.path.join(TESTS_DIR, "golden", "readme_kat.json")))
    inp = tmp_path / "input.csv"
    inp.write_text(kat["csv"])
    for metric in ("EUCLIDEAN", "TAXICAB", "COSINE"):
        out = tmp_path / f"out_{metric}"
        props = "0" if metric == "COSINE" else "10"
        print(f"[gpu] cli {metric}", flush=True)
        r = subprocess.run([os.path.join(PKG_DIR, "aknn"), str(inp), str(out), "2", "10", "10", props, metric, "0",
                            "F", "--debug"], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "has done task in:" in r.stdout
        assert out.read_bytes().hex() == kat["graphs"][metric]["result_file_hex"]
        assert (tmp_path / f"out_{metric}VALIDATION").read_text() == "Everything is OK."
        assert (tmp_path / f"out_{metric}DEBUG").read_text().startswith("NEAR: <0, ")


def test_cli_exit_codes_with_device(tmp_path):
    exe = os.path.join(PKG_DIR, "aknn")
    r = subprocess.run([exe, str(tmp_path / "missing.csv"), "o", "2", "1", "1", "1", "EUCLIDEAN", "0", "F"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 7 and "Error opening the file" in r.stdout
    bad = tmp_path / "short.csv"
    bad.write_text("4\n2\n1, 2\n3, 4\n")
    r = subprocess.run([exe, str(bad), "o", "2", "1", "1", "1", "EUCLIDEAN", "0", "F"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 8 and "Initialization failed." in r.stdout


def test_c1_end_to_end_bit_exact(aknn, oracle, datasets):
    """Config 1: N=20,000, D=32, 10 clusters, k=10, T=10, L=32, P=3, EUCLIDEAN."""
    x = datasets.gmm_c1()[0]
    g = gpu_run(aknn, x, 10, 10, 32, 3, 0)
    xn = check_graph(oracle, g, x, 10, 10, 32, 3, 0)
    check_forest(oracle, g, xn, 10, 32, 10)
    st = g.stage_times()
    assert st["total"] > 0


@pytest.mark.parametrize("n,d", [(5, 1), (479, 17), (480, 16), (481, 15), (961, 33), (1441, 64), (20000, 130),
                                 (3000, 300)])
def test_normalize_shapes_bit_exact(aknn, oracle, n, d):
    """The per-dimension sequential mean walks the rows in tiles of 480 and the
    dimensions in groups of 16: shapes on both sides of those edges.  An
    offset of 1000 makes the FP32 sum's rounding depend on the order, so a
    reordered or split sum fails."""
    rng = np.random.default_rng(n * 1000 + d)
    x = (1000.0 + rng.standard_normal((n, d)) * rng.uniform(0.01, 10.0, size=d)).astype(np.float32)
    k = min(3, n - 1)
    g = gpu_run(aknn, x, k, 1, 16, 0, 0)
    xn = oracle.normalize(x)
    assert (bits(g.normalized()) == bits(xn)).all(), "normalized points"
    if n >= 1000:
        # the data is order-sensitive: summing two halves separately rounds differently
        seq = np.zeros(d, np.float32)
        for r in x:
            seq = seq + r
        half = [np.zeros(d, np.float32), np.zeros(d, np.float32)]
        for i, r in enumerate(x):
            half[i * 2 // n] += r
        assert (half[0] + half[1] != seq).any()


@pytest.mark.parametrize("metric", [0, 1])
@pytest.mark.parametrize("props", [0, 2])
def test_small_bit_exact(aknn, oracle, datasets, metric, props):
    x = datasets.gmm_c1(3000, 24, clusters=6, seed=3)[0]
    g = gpu_run(aknn, x, 8, 4, 20, props, metric)
    check_graph(oracle, g, x, 8, 4, 20, props, metric)


@pytest.mark.parametrize("cosine_fixed", [0, 1])
@pytest.mark.parametrize("props", [0, 2])
def test_cosine_bit_exact(aknn, oracle, datasets, cosine_fixed, props):
    x = datasets.gmm_c1(2500, 20, clusters=5, seed=4)[0]
    g = gpu_run(aknn, x, 6, 3, 24, props, 2, cosine_fixed=cosine_fixed)
    check_graph(oracle, g, x, 6, 3, 24, props, 2, cosine_fixed=cosine_fixed)


@pytest.mark.parametrize("n,d,k,trees,leaf,kind", [(2500, 20, 6, 3, 24, "gmm"), (1500, 300, 7, 3, 50, "sparse"),
                                                   (3000, 40, 30, 5, 60, "gmm"), (800, 16, 12, 2, 0, "gmm")])
def test_cosine_parity_key_table(aknn, oracle, datasets, n, d, k, trees, leaf, kind):
    """COSINE parity initial pass from the per-tree key table (default) and
    from per-point row streaming: the same graph, the oracle's (D = 300 takes
    the 256-dimension chunked chains; leaf 0 gives one-point leaves whose start
    ranges are their ancestors')."""
    x = (datasets.gmm_c1(n, d, clusters=5, seed=80 + k)[0] if kind == "gmm"
         else datasets.sparse_c3(n, d, clusters=6, seed=81))
    a = gpu_run(aknn, x, k, trees, leaf, 1, 2)
    assert a.candidate_counts("cosine_table")[0] == 1.0
    b = gpu_run(aknn, x, k, trees, leaf, 1, 2, cosine_table=0)
    assert b.candidate_counts("cosine_table")[0] == 0.0
    _same_graph(a, b)
    check_graph(oracle, a, x, k, trees, leaf, 1, 2)


@pytest.mark.parametrize("metric", [0, 1, 2])
def test_wide_rows_bit_exact(aknn, oracle, datasets, metric):
    """D=300 > 256: the normalization quirk and 256-dimension chunked sums."""
    x = datasets.sparse_c3(1200, 300, clusters=8, seed=6)
    g = gpu_run(aknn, x, 5, 2, 40, 1, metric)
    check_graph(oracle, g, x, 5, 2, 40, 1, metric)


@pytest.mark.parametrize("leaf,k", [(0, 3), (1, 4), (7, 12), (5000, 9)])
def test_forest_edge_leaf_sizes(aknn, oracle, datasets, leaf, k):
    x = datasets.uniform_c2(1500, 12, seed=leaf + 1)
    g = gpu_run(aknn, x, k, 3, leaf, 1, 0)
    xn = check_graph(oracle, g, x, k, 3, leaf, 1, 0)
    check_forest(oracle, g, xn, 3, leaf, k)


@pytest.mark.parametrize("seed,order", [(2, 0), (1, 1), (987654321, 0)])
def test_rng_options(aknn, oracle, datasets, seed, order):
    x = datasets.gmm_c1(2000, 16, clusters=4, seed=9)[0]
    g = gpu_run(aknn, x, 6, 3, 16, 1, 0, seed=seed, rng_order=order)
    xn = check_graph(oracle, g, x, 6, 3, 16, 1, 0, seed=seed, order=order)
    check_forest(oracle, g, xn, 3, 16, 6, seed=seed, order=order)


def test_level_batch_does_not_change_results(aknn, datasets):
    x = datasets.gmm_c1(3000, 16, clusters=5, seed=2)[0]
    a = gpu_run(aknn, x, 8, 3, 8, 1, 0, level_batch=1).graph()
    b = gpu_run(aknn, x, 8, 3, 8, 1, 0, level_batch=64).graph()
    assert (a[0] == b[0]).all() and (bits(a[1]) == bits(b[1])).all()


def test_recalculate_is_identical(aknn, datasets):
    x = datasets.gmm_c1(4000, 16, clusters=5, seed=8)[0]
    g = gpu_run(aknn, x, 10, 4, 16, 2, 0)
    a = g.graph()
    g.calculate()
    b = g.graph()
    assert (a[0] == b[0]).all() and (bits(a[1]) == bits(b[1])).all()


def test_k_equals_n_minus_one(aknn, oracle, datasets):
    x = datasets.uniform_c2(40, 5, seed=3)
    g = gpu_run(aknn, x, 39, 2, 4, 1, 0)
    check_graph(oracle, g, x, 39, 2, 4, 1, 0)


def test_rejects_k_at_least_n(aknn, datasets):
    x = datasets.uniform_c2(10, 3)
    g = aknn.Graph()
    with pytest.raises(aknn.AknnError, match="k\\+1"):
        g.initialize(x, 10, 2, 4, 1, 0)


@pytest.mark.parametrize("metric,fixed", [(0, 0), (1, 0), (2, 1)])
def test_skip_settled_changes_nothing(aknn, oracle, datasets, metric, fixed):
    """Propagation's skip of already-examined neighbours gives the same graph (and the oracle's)."""
    x = datasets.gmm_c1(4000, 16, clusters=6, seed=21)[0]
    a = gpu_run(aknn, x, 8, 2, 24, 5, metric, cosine_fixed=fixed, skip_settled=1)
    b = gpu_run(aknn, x, 8, 2, 24, 5, metric, cosine_fixed=fixed, skip_settled=0)
    ga, gb = a.graph(), b.graph()
    assert (ga[0] == gb[0]).all() and (bits(ga[1]) == bits(gb[1])).all()
    sa, sb = a.kernel_stats("propagate"), b.kernel_stats("propagate")
    assert sa["units"] < sb["units"]  # fewer candidate rows examined
    check_graph(oracle, a, x, 8, 2, 24, 5, metric, cosine_fixed=fixed)


def _same_graph(a, b):
    ga, gb = a.graph(), b.graph()
    assert (ga[0] == gb[0]).all() and (bits(ga[1]) == bits(gb[1])).all()


@pytest.mark.parametrize("metric,fixed,d", [(0, 0, 40), (1, 0, 40), (2, 1, 40), (2, 0, 40), (0, 0, 128), (1, 0, 96),
                                            (2, 1, 128)])
def test_sketch_changes_nothing(aknn, oracle, datasets, metric, fixed, d):
    """The FP16 sketch screen of propagation gives the same graph (and the oracle's), re-scoring fewer candidates."""
    x = datasets.gmm_c1(4000, d, clusters=6, seed=31)[0]
    a = gpu_run(aknn, x, 8, 2, 24, 4, metric, cosine_fixed=fixed, sketch=1)
    b = gpu_run(aknn, x, 8, 2, 24, 4, metric, cosine_fixed=fixed, sketch=0)
    _same_graph(a, b)
    ex, rs = a.candidate_counts()
    ex0, rs0 = b.candidate_counts()
    print(f"[gpu] examined {ex:.0f} rescored {rs:.0f} (plain {rs0:.0f})", flush=True)
    assert ex0 == rs0
    # COSINE parity screens its candidates' own first k (the flip makes the
    # row's k-th value no threshold); the others screen against the row
    assert ex == ex0 and 0 < rs < ex
    check_graph(oracle, a, x, 8, 2, 24, 4, metric, cosine_fixed=fixed)


@pytest.mark.parametrize("flip8", [1, 0])
@pytest.mark.parametrize("n,d,k,props", [(5000, 96, 30, 3), (3000, 300, 20, 2), (2000, 24, 12, 4)])
def test_cosine_parity_screen(aknn, oracle, datasets, n, d, k, props, flip8):
    """COSINE parity rounds with the sketch screen: the candidates' own first
    k bounded by the k-th smallest sketch upper bound (the int8 sketch's with
    flip_int8, the FP16 one's without); the oracle's graph, and that of the
    unscreened rounds, with fewer candidates keyed exactly."""
    x = datasets.sphere_c4(n, d, seed=7) if d == 96 else datasets.gmm_c1(n, d, clusters=8, seed=7)[0]
    print(f"[gpu] COSINE parity screen n={n} d={d} k={k} rounds={props} flip_int8={flip8}", flush=True)
    a = gpu_run(aknn, x, k, 3, 2 * k, props, 2, sketch=1, flip_int8=flip8)
    b = gpu_run(aknn, x, k, 3, 2 * k, props, 2, sketch=0)
    _same_graph(a, b)
    ex, rs = a.candidate_counts()
    assert 0 < rs < ex
    check_graph(oracle, a, x, k, 3, 2 * k, props, 2)


def _adversarial(n=3000, d=300, seed=5):
    """Exact duplicates (zero distances, id ties), values far beyond FP16 in the
    dimensions the normalization leaves unscaled (d >= 256), and rows in the
    FP16 subnormal range."""
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, d)).astype(np.float32)
    x[100:130] = x[100]                       # 30 identical points (fewer than the leaf bound: a
    #                                           bucket of identical points can never split)
    x[400:700:3] = x[400] + np.float32(1e-6) * rng.normal(size=(100, d)).astype(np.float32)
    x[700:760, 280:] *= np.float32(1e6)       # beyond FP16 after normalization (unscaled tail)
    x[800:900] = x[0] + np.float32(3e-5) * rng.normal(size=(100, d)).astype(np.float32)
    return x


@pytest.mark.parametrize("metric,fixed", [(0, 0), (1, 0), (2, 1), (2, 0)])
def test_sketch_adversarial_inputs(aknn, oracle, metric, fixed):
    x = _adversarial()
    a = gpu_run(aknn, x, 10, 2, 40, 3, metric, cosine_fixed=fixed, sketch=1)
    b = gpu_run(aknn, x, 10, 2, 40, 3, metric, cosine_fixed=fixed, sketch=0)
    _same_graph(a, b)
    xn = check_graph(oracle, a, x, 10, 2, 40, 3, metric, cosine_fixed=fixed)
    check_forest(oracle, a, xn, 2, 40, 10)


def test_sketch_c1_bit_exact(aknn, oracle, datasets):
    x = datasets.gmm_c1()[0]
    g = gpu_run(aknn, x, 10, 10, 32, 3, 0, sketch=1)
    xn = check_graph(oracle, g, x, 10, 10, 32, 3, 0)
    check_forest(oracle, g, xn, 10, 32, 10)


@pytest.mark.parametrize("n,d,k,trees,leaf,seed", [(4000, 40, 8, 4, 24, 41), (3000, 128, 20, 3, 100, 42),
                                                   (1500, 300, 10, 3, 60, 43)])
def test_leaf_mfma_changes_nothing(aknn, oracle, datasets, n, d, k, trees, leaf, seed):
    """Matrix-core screening of the leaf pairs (EUCLIDEAN) gives the plain graph and the oracle's."""
    x = datasets.gmm_c1(n, d, clusters=6, seed=seed)[0]
    a = gpu_run(aknn, x, k, trees, leaf, 0, 0, leaf_mfma=1)
    b = gpu_run(aknn, x, k, trees, leaf, 0, 0, leaf_mfma=0)
    _same_graph(a, b)
    scr, res = a.candidate_counts("leaf_knn")
    print(f"[gpu] leaf pairs screened {scr:.0f} rescored {res:.0f}", flush=True)
    assert scr > 0 and res < scr
    check_graph(oracle, a, x, k, trees, leaf, 0, 0)


@pytest.mark.parametrize("k", [23, 40, 64])
@pytest.mark.parametrize("leaf_mfma", [0, 1])
def test_leaf_list_sizes(aknn, oracle, datasets, k, leaf_mfma):
    """Owner lists of 24 and 64 slots holding 23, 40 and 64 pairs: the row sits
    in the top slots of the list with empty slots below it (no propagation
    round: the leaf kernels alone)."""
    x = datasets.gmm_c1(3000, 40, clusters=5, seed=50 + k)[0]
    g = gpu_run(aknn, x, k, 3, 80, 0, 0, leaf_mfma=leaf_mfma)
    check_graph(oracle, g, x, k, 3, 80, 0, 0)


@pytest.mark.parametrize("k,metric,opts", [(48, 0, {}), (48, 0, {"sketch": 1}), (64, 0, {"sketch": 1, "leaf_mfma": 1}),
                                           (64, 1, {}), (64, 2, {}), (64, 2, {"cosine_fixed": 1, "sketch": 1})])
def test_large_k_propagation(aknn, oracle, datasets, k, metric, opts):
    """Propagation rounds at k = 48 and 64 (the reference propagates any k,
    .cu:2768, :2797-2805): each wave's hash set and candidate list
    (2(k^2+k+1) and k^2 slots) no longer fit four waves per block, so the
    launch takes fewer waves per block."""
    x = datasets.gmm_c1(2500, 24, clusters=5, seed=60 + k + metric)[0]
    g = gpu_run(aknn, x, k, 3, 80, 2, metric, **opts)
    check_graph(oracle, g, x, k, 3, 80, 2, metric, cosine_fixed=opts.get("cosine_fixed", 0))


@pytest.mark.parametrize("metric,opts", [(0, {"sketch": 1, "leaf_mfma": 1}), (1, {"sketch": 1}),
                                         (2, {"sketch": 1, "cosine_fixed": 1})])
def test_wide_rows_screened_propagation(aknn, oracle, datasets, metric, opts):
    """D = 3000 with the FP16 sketch: one wave's x_i, sketch row and two
    staged rows take 48 KB of LDS, so the screened rounds run two waves per
    block."""
    x = datasets.gmm_c1(1200, 3000, clusters=4, seed=71 + metric)[0]
    g = gpu_run(aknn, x, 20, 2, 60, 2, metric, **opts)
    check_graph(oracle, g, x, 20, 2, 60, 2, metric, cosine_fixed=opts.get("cosine_fixed", 0))


@pytest.mark.parametrize("n,d,k,trees,leaf,seed", [(4000, 40, 8, 4, 24, 45), (2500, 96, 30, 3, 100, 46),
                                                   (1500, 300, 10, 3, 60, 47)])
def test_leaf_mfma_cosine_fixed(aknn, oracle, datasets, n, d, k, trees, leaf, seed):
    """Matrix-core screening of COSINE leaf pairs (carried distances): the plain graph and the oracle's."""
    x = datasets.gmm_c1(n, d, clusters=6, seed=seed)[0]
    a = gpu_run(aknn, x, k, trees, leaf, 0, 2, cosine_fixed=1, leaf_mfma=1)
    b = gpu_run(aknn, x, k, trees, leaf, 0, 2, cosine_fixed=1, leaf_mfma=0)
    _same_graph(a, b)
    scr, res = a.candidate_counts("leaf_knn")
    print(f"[gpu] cosine leaf pairs screened {scr:.0f} rescored {res:.0f}", flush=True)
    assert scr > 0 and res < scr
    check_graph(oracle, a, x, k, trees, leaf, 0, 2, cosine_fixed=1)


def test_leaf_mfma_uniform_and_adversarial(aknn, oracle, datasets):
    x = datasets.uniform_c2(5000, 128, seed=44)
    a = gpu_run(aknn, x, 20, 4, 256, 1, 0, leaf_mfma=1, sketch=1)
    check_graph(oracle, a, x, 20, 4, 256, 1, 0)
    x = _adversarial()
    a = gpu_run(aknn, x, 10, 3, 40, 0, 0, leaf_mfma=1)
    check_graph(oracle, a, x, 10, 3, 40, 0, 0)
    a = gpu_run(aknn, x, 10, 3, 40, 1, 2, leaf_mfma=1, sketch=1, cosine_fixed=1)
    check_graph(oracle, a, x, 10, 3, 40, 1, 2, cosine_fixed=1)


def test_chosen_form_c1_bit_exact(aknn, oracle, datasets):
    x = datasets.gmm_c1()[0]
    g = gpu_run(aknn, x, 10, 10, 32, 3, 0, sketch=1, leaf_mfma=1)
    check_graph(oracle, g, x, 10, 10, 32, 3, 0)


@pytest.mark.parametrize("d,metric", [(20, 0), (128, 0), (300, 1), (33, 2)])
def test_forest_sketch_same_trees(aknn, oracle, datasets, d, metric):
    """The sketch-decided side tests give the oracle's trees (leaf order, start ranges, rand() use)."""
    x = datasets.gmm_c1(3000, d, clusters=5, seed=50 + d)[0]
    g = gpu_run(aknn, x, 8, 3, 20, 1, metric, sketch=1, cosine_fixed=1)
    ex, rs = g.candidate_counts("forest")
    print(f"[gpu] side tests {ex:.0f} exact {rs:.0f}", flush=True)
    assert ex > 0 and rs < ex
    xn = check_graph(oracle, g, x, 8, 3, 20, 1, metric, cosine_fixed=1)
    check_forest(oracle, g, xn, 3, 20, 8)


@pytest.mark.parametrize("kind,d,leaf", [("gmm", 128, 60), ("uniform", 128, 256), ("gmm", 33, 16),
                                          ("gmm", 300, 100), ("adversarial", 300, 40)])
def test_forest_int8_same_trees(aknn, oracle, datasets, kind, d, leaf):
    """The int8 side screen of staged buckets (option forest_int8, the
    default) against the FP16 one and the oracle: the same trees.  The int8
    bound leaves more points to the exact chain."""
    if kind == "gmm":
        x = datasets.gmm_c1(4000, d, clusters=6, seed=70 + d)[0]
    elif kind == "uniform":
        x = datasets.uniform_c2(4000, d, seed=71)
    else:
        x = _adversarial(n=3000, d=d)
    counts = []
    for f8 in (1, 0):
        g = gpu_run(aknn, x, 10, 3, leaf, 0, 0, sketch=1, forest_int8=f8)
        ex, rs = g.candidate_counts("forest")
        print(f"[gpu] forest_int8={f8}: side tests {ex:.0f} exact {rs:.0f}", flush=True)
        counts.append((ex, rs))
        xn = g.normalized()
        check_forest(oracle, g, xn, 3, leaf, 10)
        g.close()
    assert counts[0][0] == counts[1][0] and counts[0][1] >= counts[1][1] and counts[0][1] < counts[0][0]


@pytest.mark.parametrize("int8", [0, 1])
def test_sketch_precisions_same_graph(aknn, oracle, datasets, int8):
    """EUCLIDEAN propagation screened by the int8 or the FP16 sketch: the plain graph and the oracle's."""
    x = datasets.uniform_c2(6000, 64, seed=60)
    a = gpu_run(aknn, x, 12, 3, 40, 4, 0, sketch=1, sketch_int8=int8)
    b = gpu_run(aknn, x, 12, 3, 40, 4, 0)
    _same_graph(a, b)
    ex, rs = a.candidate_counts()
    print(f"[gpu] int8={int8} examined {ex:.0f} rescored {rs:.0f}", flush=True)
    assert rs < ex
    check_graph(oracle, a, x, 12, 3, 40, 4, 0)


@pytest.mark.parametrize("metric,d", [(0, 24), (1, 24), (2, 24), (0, 300), (2, 130)])
def test_exact_knn_matches_oracle(aknn, oracle, datasets, metric, d):
    """FP64 exact k-NN on the GPU: k-th distances and every returned distance bit-equal to the oracle's FP64."""
    x = datasets.gmm_c1(3000, d, clusters=5, seed=70 + d)[0]
    g = gpu_run(aknn, x, 10, 2, 32, 1, metric, cosine_fixed=1)
    xn = g.normalized()
    q = np.random.default_rng(d).choice(len(x), 150, replace=False)
    ids, dist = g.exact_knn(q, 10)
    kth = oracle.exact_kth(xn, 10, metric, q)
    assert (dist[:, -1].view(np.uint64) == kth.view(np.uint64)).all()
    lib = oracle.lib()
    for r, qq in enumerate(q[:40]):
        assert qq not in ids[r]
        ref = [lib.orc_distance64(np.ascontiguousarray(xn[qq]), np.ascontiguousarray(xn[c]), d, metric)
               for c in ids[r]]
        assert (np.array(ref).view(np.uint64) == dist[r].view(np.uint64)).all()
        order = sorted(zip(dist[r].tolist(), ids[r].tolist()))
        assert [c for _, c in order] == ids[r].tolist()
    rec = g.recall(q)
    gi, _ = g.graph()
    assert rec == oracle.recall(xn, 10, metric, q, kth, gi)
