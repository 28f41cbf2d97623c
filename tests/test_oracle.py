"""Pins the CPU oracle before it is trusted as the checker.

Every check is independent of the oracle's own code: libc's rand(), exact
rational arithmetic (tests/golden/make_kat.py), numpy's IEEE float32 ops,
textbook brute force, and invariants a wrong index or sign would break.
"""
import ctypes
import importlib.util
import json
import os
import struct

import numpy as np
import pytest

from conftest import TESTS_DIR

GOLDEN = os.path.join(TESTS_DIR, "golden")


def _kat():
    with open(os.path.join(GOLDEN, "readme_kat.json")) as f:
        return json.load(f)


def _exact():
    spec = importlib.util.spec_from_file_location("make_kat", os.path.join(GOLDEN, "make_kat.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def bits32(v):
    return int(np.float32(v).view(np.uint32))


# --- rand() -----------------------------------------------------------------

@pytest.mark.parametrize("seed", [1, 2, 77, 12345, 0xDEADBEEF])
def test_rng_matches_libc_rand(oracle, seed):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(ctypes.c_uint(seed))
    want = [libc.rand() for _ in range(3000)]
    g = oracle.rng(seed)
    got = [oracle.lib().orc_rng_next(ctypes.byref(g)) for _ in range(3000)]
    assert got == want


def test_pick_is_the_double_expression(oracle):
    rng = np.random.default_rng(0)
    for r in rng.integers(0, 2**31 - 1, 2000).tolist() + [0, 1, 2**31 - 2, 2**31 - 1]:
        for c in (1, 2, 3, 7, 1000, 2**20 + 3):
            want = int((r / 2147483647.0) * 2147483647.0) % c
            assert oracle.lib().orc_pick(r, c) == want


# --- README known answers (exact arithmetic) ----------------------------------

def test_readme_normalization(oracle):
    kat = _kat()
    raw = np.array(kat["raw_bits"], np.uint32).view(np.float32)
    x = oracle.normalize(raw)
    assert (bits(x) == np.array(kat["normalized_bits"], np.uint32)).all()


@pytest.mark.parametrize("metric", ["EUCLIDEAN", "TAXICAB", "COSINE"])
def test_readme_graph_and_file(oracle, metric, tmp_path):
    kat = _kat()
    raw = np.array(kat["raw_bits"], np.uint32).view(np.float32)
    props = 0 if metric == "COSINE" else 10  # COSINE's stub survives; the reference reads out of bounds
    x, ids, vals, _ = oracle.run(raw, kat["k"], 10, 10, props, oracle.METRICS[metric])
    g = kat["graphs"][metric]
    assert ids.tolist() == g["ids"]
    assert (bits(vals) == np.array(g["val_bits"], np.uint32)).all()
    out = str(tmp_path / "res")
    oracle.write_result(out, ids, vals, oracle.METRICS[metric])
    assert open(out, "rb").read().hex() == g["result_file_hex"]


def test_readme_euclidean_table_from_survey():
    # hand-derived in the survey: point 1's order [2, 0] exists only because of
    # the sequential FP32 mean
    g = _kat()["graphs"]["EUCLIDEAN"]
    assert g["ids"] == [[1, 2], [2, 0], [1, 0]]
    assert g["val_bits"] == [[0x3FB504F3, 0x403504F3], [0x3FB504F2, 0x3FB504F3], [0x3FB504F2, 0x403504F3]]


def test_distance_kats(oracle):
    for dk in _kat()["distance_kats"]:
        a = np.array(dk["a_bits"], np.uint32).view(np.float32)
        b = np.array(dk["b_bits"], np.uint32).view(np.float32)
        for m, v in dk["bits"].items():
            assert bits32(oracle.distance(a, b, oracle.METRICS[m])) == v, m
    # (0,0)-(3,4): 5 and sqrt(7); orthogonal unit vectors: key 1
    assert oracle.distance([0, 0], [3, 4], 0) == np.float32(5)
    assert oracle.distance([0, 0], [3, 4], 1) == np.sqrt(np.float32(7))
    assert oracle.distance([1, 0], [0, 1], 2) == np.float32(1)


@pytest.mark.parametrize("metric", ["EUCLIDEAN", "TAXICAB", "COSINE"])
def test_distance_random_exact(oracle, metric):
    ex = _exact()
    rng = np.random.default_rng(11)
    for d in (1, 5, 31, 256, 300):
        for _ in range(3):
            a = rng.normal(0, 1, d).astype(np.float32)
            b = rng.normal(0, 1, d).astype(np.float32)
            want = ex.distance([ex.F(float(v)) for v in a], [ex.F(float(v)) for v in b], metric)
            assert bits32(oracle.distance(a, b, oracle.METRICS[metric])) == ex.f32_bits(want)


def test_distance_symmetric(oracle):
    rng = np.random.default_rng(3)
    for m in (0, 1, 2):
        for d in (3, 64, 257):
            a, b = rng.normal(size=(2, d)).astype(np.float32)
            assert bits32(oracle.distance(a, b, m)) == bits32(oracle.distance(b, a, m))


# --- normalization: independent numpy restatement ------------------------------

@pytest.mark.parametrize("shape", [(500, 7), (300, 300), (1000, 32)])
def test_normalize_vs_numpy(oracle, shape):
    rng = np.random.default_rng(5)
    x = (rng.normal(3, 2, shape) * 10).astype(np.float32)
    got = oracle.normalize(x)
    n, d = shape
    want = x.copy()
    for j in range(d):
        m = np.cumsum(want[:, j], dtype=np.float32)[-1] / np.float32(n)  # sequential float32 sum
        want[:, j] = want[:, j] - np.float32(m)
    dm = min(d, 256)
    g = np.abs(want[:, :dm]).max()
    want[:, :dm] = want[:, :dm] / g
    assert (bits(got) == bits(want)).all()
    if d > 256:  # dimensions >= 256 are centred but not scaled (.cu:57-84)
        assert np.abs(got[:, 256:]).max() > 1.0


# --- forest invariants -----------------------------------------------------------

def _forest(oracle, x, trees, leaf, k, seed=1, order=0):
    return oracle.build_forest(x, trees, leaf, k, seed=seed, rng_order=order)


@pytest.mark.parametrize("leaf,k", [(8, 3), (32, 10), (1, 2), (0, 2), (50, 49)])
def test_forest_invariants(oracle, leaf, k):
    rng = np.random.default_rng(leaf + 7 * k)
    n = 1500
    x = oracle.normalize(rng.normal(0, 1, (n, 6)).astype(np.float32))
    perm, rs, rc, info = _forest(oracle, x, 3, leaf, k)
    for t in range(3):
        assert sorted(perm[t].tolist()) == list(range(n))
        pos = np.empty(n, np.int64)
        pos[perm[t]] = np.arange(n)
        # every point lies inside its start range; ranges hold >= k+1 points
        assert ((rs[t] <= pos) & (pos < rs[t] + rc[t])).all()
        assert (rc[t] >= k + 1).all()
        assert ((rs[t] >= 0) & (rs[t] + rc[t] <= n)).all()
        # ranges nest or are disjoint (they are subtrees of one tree)
        uniq = sorted(set(zip(rs[t].tolist(), rc[t].tolist())))
        for a, (s1, c1) in enumerate(uniq):
            for s2, c2 in uniq[a + 1:]:
                disjoint = s2 >= s1 + c1 or s1 >= s2 + c2
                nested = (s1 <= s2 and s2 + c2 <= s1 + c1) or (s2 <= s1 and s1 + c1 <= s2 + c2)
                assert disjoint or nested
        it, nodes, active = info[t]
        assert nodes % 2 == 1 and it >= 1


def test_forest_single_leaf(oracle):
    rng = np.random.default_rng(1)
    n = 40
    x = oracle.normalize(rng.normal(size=(n, 3)).astype(np.float32))
    perm, rs, rc, info = _forest(oracle, x, 2, 64, 5)
    assert (perm == np.arange(n)).all()
    assert (rs == 0).all() and (rc == n).all()
    assert [i[0] for i in info] == [1, 1]  # one iteration, nothing split


def test_forest_depends_on_rng(oracle):
    rng = np.random.default_rng(2)
    x = oracle.normalize(rng.normal(size=(800, 5)).astype(np.float32))
    a = _forest(oracle, x, 2, 16, 4, seed=1)[0]
    b = _forest(oracle, x, 2, 16, 4, seed=2)[0]
    c = _forest(oracle, x, 2, 16, 4, seed=1, order=1)[0]
    assert (a != b).any() and (a != c).any()
    assert (a == _forest(oracle, x, 2, 16, 4, seed=1)[0]).all()


def test_forest_split_rule_by_hand(oracle):
    # 1-D, 8 points, leaf 7: exactly one split of the root.  Side = sign of
    # x * (x_p2 - x_p1) with pivots from the first two rand() draws, p1 forced
    # LEFT and p2 RIGHT, then a stable partition.
    g = oracle.rng(1)
    a, b = oracle.lib().orc_rng_next(ctypes.byref(g)), oracle.lib().orc_rng_next(ctypes.byref(g))
    vals = np.array([-3, -1, 2, 5, -2, 4, 1, -4], np.float32)
    n = len(vals)
    q1, q2 = a % n, b % n
    if q1 == q2:
        q2 = (q1 + 1) % n
    perm, rs, rc, info = oracle.build_forest(vals.reshape(-1, 1), 1, 7, 1)
    s = vals * (vals[q2] - vals[q1])
    right = ~(s < 0)
    right[q1] = False
    right[q2] = True
    want = [i for i in range(n) if not right[i]] + [i for i in range(n) if right[i]]
    assert perm[0].tolist() == want
    assert info[0][0] == 2  # one splitting iteration + the final one


# --- initial pass and propagation against brute force --------------------------

def _topk(pool, k):
    return sorted(pool, key=lambda e: (e[0], e[1]))[:k]


@pytest.mark.parametrize("metric", [0, 1])
def test_initial_single_leaf_is_brute_force(oracle, metric):
    rng = np.random.default_rng(9)
    n, k = 300, 7
    x = oracle.normalize(rng.normal(size=(n, 10)).astype(np.float32))
    perm, rs, rc, _ = _forest(oracle, x, 1, n, k)  # L >= N: every candidate set is all others
    ids, vals = oracle.initial_knn(x, k, metric, perm, rs, rc)
    for i in range(0, n, 17):
        pool = [(float(oracle.distance(x[i], x[j], metric)), j) for j in range(n) if j != i]
        want = _topk(pool, k)
        assert ids[i].tolist() == [j for _, j in want]
        assert vals[i].tolist() == [np.float32(v) for v, _ in want]


def test_initial_cosine_chunks_flip(oracle):
    # k=3, N=6: cap = min(12, 6) - 3 = 3, two chunks; replay the chunk loop in Python
    rng = np.random.default_rng(4)
    n, k = 6, 3
    x = oracle.normalize(rng.normal(size=(n, 4)).astype(np.float32))
    perm, rs, rc, _ = _forest(oracle, x, 1, n, k)
    ids, vals = oracle.initial_knn(x, k, 2, perm, rs, rc)
    for i in range(n):
        cand = [j for j in range(n) if j != i]
        lst = [(100000.0, -1)] * k
        for c0 in range(0, len(cand), 3):
            pool = lst + [(float(oracle.distance(x[i], x[j], 2)), j) for j in cand[c0:c0 + 3]]
            lst = [(float(np.float32(1.0 - v)), j) for v, j in _topk(pool, k)]
        lst = _topk(lst, k)
        assert ids[i].tolist() == [j for _, j in lst]
        assert vals[i].tolist() == [np.float32(v) for v, _ in lst]


@pytest.mark.parametrize("metric", [0, 1, 2])
def test_propagation_brute_force(oracle, metric):
    rng = np.random.default_rng(21 + metric)
    n, k = 120, 5
    x = oracle.normalize(rng.normal(size=(n, 8)).astype(np.float32))
    # a random graph with valid rows (no self, distinct), values = real keys, sorted
    ids = np.zeros((n, k), np.int32)
    vals = np.zeros((n, k), np.float32)
    for i in range(n):
        c = rng.choice([j for j in range(n) if j != i], k, replace=False)
        row = _topk([(float(oracle.distance(x[i], x[j], metric)), int(j)) for j in c], k)
        ids[i] = [j for _, j in row]
        vals[i] = [v for v, _ in row]
    io, vo = oracle.propagate(x, ids, vals, metric)
    for i in range(n):
        own = set(ids[i].tolist())
        cand = sorted({int(c) for j in ids[i] for c in ids[j]} - {i} - own)
        if not cand:
            want = list(zip(vals[i].tolist(), ids[i].tolist()))
        else:
            pool = list(zip(vals[i].tolist(), ids[i].tolist()))
            pool += [(float(oracle.distance(x[i], x[j], metric)), j) for j in cand]
            want = _topk(pool, k)
            if metric == 2:
                want = _topk([(float(np.float32(1.0 - v)), j) for v, j in want], k)
        assert io[i].tolist() == [j for _, j in want]
        assert vo[i].tolist() == [np.float32(v) for v, _ in want]


@pytest.mark.parametrize("metric", [0, 1])
def test_propagation_fixed_point(oracle, metric):
    # the exact FP32 k-NN graph does not change under a round (E and T)
    rng = np.random.default_rng(8)
    n, k = 200, 6
    x = oracle.normalize(rng.normal(size=(n, 5)).astype(np.float32))
    perm, rs, rc, _ = _forest(oracle, x, 1, n, k)
    ids, vals = oracle.initial_knn(x, k, metric, perm, rs, rc)
    io, vo = oracle.propagate(x, ids, vals, metric)
    assert (io == ids).all() and (bits(vo) == bits(vals)).all()


def test_cosine_parity_breaks_fixed_point(oracle):
    rng = np.random.default_rng(8)
    n, k = 200, 6
    x = oracle.normalize(rng.normal(size=(n, 5)).astype(np.float32))
    perm, rs, rc, _ = _forest(oracle, x, 1, n, k)
    ids, vals = oracle.initial_knn(x, k, 2, perm, rs, rc, cosine_fixed=1)
    io, vo = oracle.propagate(x, ids, vals, 2, cosine_fixed=1)
    assert (io == ids).all()
    io2, vo2 = oracle.propagate(x, ids, vals, 2, cosine_fixed=0)
    assert (bits(vo2) != bits(vals)).any()


# --- whole pipeline, recall, writer ---------------------------------------------

def test_pipeline_deterministic_and_valid(oracle, datasets):
    x = datasets.gmm_c1(2000, 16)[0]
    a = oracle.run(x, 5, 4, 16, 2, 0)
    b = oracle.run(x, 5, 4, 16, 2, 0)
    assert (a[1] == b[1]).all() and (bits(a[2]) == bits(b[2])).all()
    ids = a[1]
    n = len(x)
    assert ((ids >= 0) & (ids < n)).all()
    assert (ids != np.arange(n)[:, None]).all()
    assert all(len(set(r)) == len(r) for r in ids.tolist())
    assert (np.diff(a[2], axis=1) >= 0).all()


def test_exact_kth_and_recall_vs_numpy(oracle):
    rng = np.random.default_rng(12)
    n, d, k = 400, 6, 5
    x = rng.normal(size=(n, d)).astype(np.float32)
    q = np.array([0, 5, 77, 399], np.int64)
    kth = oracle.exact_kth(x, k, 0, q)
    x64 = x.astype(np.float64)
    for qi, i in enumerate(q):
        dd = ((x64 - x64[i]) ** 2).sum(1)
        dd[i] = np.inf
        assert np.isclose(np.sort(dd)[k - 1], kth[qi], rtol=1e-12)
    ids = np.argsort(((x64[:, None] - x64[None]) ** 2).sum(-1) + np.diag(np.full(n, np.inf)), 1)[:, :k]
    assert oracle.recall(x, k, 0, q, kth, ids.astype(np.int32)) == 1.0


def test_writer_layout(oracle, tmp_path):
    rng = np.random.default_rng(1)
    n, k = 5, 3
    ids = rng.integers(0, 100, (n, k)).astype(np.int32)
    vals = rng.random((n, k)).astype(np.float32)
    vals[0, 1] = vals[0, 2]
    for metric in (0, 2):
        p = str(tmp_path / f"r{metric}")
        oracle.write_result(p, ids, vals, metric)
        blob = open(p, "rb").read()
        head = b"%d;%d;8\n" % (n, k)
        assert blob.startswith(head)
        assert blob[len(head):len(head) + 8] == bytes([4, 3, 2, 1, 0, 0, 0, 0])
        body = blob[len(head) + 8:]
        assert len(body) == n * k * 8
        for i in range(n):
            row = sorted(zip(vals[i].tolist(), ids[i].tolist()),
                         key=(lambda e: (-e[0], e[1])) if metric == 2 else (lambda e: e))
            for j, (v, c) in enumerate(row):
                off = (i * k + j) * 8
                assert struct.unpack("<if", body[off:off + 8]) == (c, np.float32(v))
