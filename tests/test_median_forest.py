"""The median-split forest: the oracle pinned by an independent restatement,
and (marked gpu) the HIP build against the oracle bit for bit.

The rule is this rebuild's own (the reference splits by the sign of
x . (x_p2 - x_p1), euclideandistancematrixgpu.cu:2138-2181), so there is no
reference output to match; the oracle's median tree is pinned here by
  * a pure-Python restatement of the directions (splitmix64 + Box-Muller) and
    of the split rule (sample positions, pivot index, tie order, stable
    partition, bin bits, position fallback), compared exactly;
  * what the rule guarantees: a permutation, leaves of 1..leaf points, exact
    halving when every point is sampled (c <= 511), both sides non-empty even
    for identical points, routed trees keeping bins apart down to the leaves.
"""
import math
import struct

import numpy as np
import pytest

M64 = (1 << 64) - 1


def py_directions(seed, tree, count, d):
    s = ((seed << 32) ^ (((tree + 1) * 0xD1B54A32D192ED03) & M64)) & M64
    out = []

    def nxt():
        nonlocal s
        s = (s + 0x9E3779B97F4A7C15) & M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)

    while len(out) < count * d:
        u1 = ((nxt() >> 11) + 0.5) * 2.0 ** -53
        u2 = ((nxt() >> 11) + 0.5) * 2.0 ** -53
        r, th = math.sqrt(-2.0 * math.log(u1)), 6.283185307179586 * u2
        out.append(r * math.cos(th))
        if len(out) < count * d:
            out.append(r * math.sin(th))
    return np.array(out, np.float64).astype(np.float32).reshape(count, d)


def ordkey(f):
    b = struct.unpack("<I", struct.pack("<f", float(f)))[0]
    return (~b) & 0xFFFFFFFF if b & 0x80000000 else b | 0x80000000


def fmaf(a, b, c):
    """float32 fused multiply-add, rounded once.  The product of two floats is
    exact in double; the double sum t is off the exact sum by e (TwoSum), and
    rounding t to float32 differs from rounding the exact sum only when t sits
    exactly on a float32 midpoint, where e decides."""
    p = float(a) * float(b)
    c = float(c)
    t = p + c
    bp = t - p
    e = (p - (t - bp)) + (c - bp)
    r = np.float32(t)
    if e != 0.0 and float(r) != t:
        other = np.nextafter(r, np.float32(np.inf) if t > float(r) else np.float32(-np.inf))
        if (float(r) + float(other)) / 2.0 == t:
            up = e > 0
            r = max(r, other) if up else min(r, other)
    return np.float32(r)


def fma_chain(row, r):
    acc = np.float32(0.0)
    for a, b in zip(row, r):
        acc = fmaf(a, b, acc)
    return acc


def py_median_tree(x, leaf, k, dirs, bins=None, bin_bits=0):
    n = len(x)
    perm = list(range(n))
    rs = [0] * n
    rc = [0] * n
    leaf = max(leaf, 1)
    bb = bin_bits if bins is not None else 0

    def bucket(s, c, lv, ps, pc):
        r0, c0 = (s, c) if c >= k + 1 else (ps, pc)
        if c <= leaf:
            for p in range(s, s + c):
                rs[perm[p]], rc[perm[p]] = r0, c0
            return
        if lv < bb:
            side = [(bins[perm[p]] >> (bb - 1 - lv)) & 1 for p in range(s, s + c)]
        elif lv - bb >= len(dirs):
            side = [1 if rel >= c // 2 else 0 for rel in range(c)]
        else:
            r = dirs[lv - bb]
            proj = [fma_chain(x[perm[p]], r) for p in range(s, s + c)]
            m = min(c, 511)
            smp = sorted((ordkey(proj[q * c // m]), q * c // m) for q in range(m))
            piv = smp[m // 2]
            side = [1 if (ordkey(proj[rel]), rel) >= piv else 0 for rel in range(c)]
        seg = perm[s:s + c]
        left = [e for e, sd in zip(seg, side) if sd == 0]
        right = [e for e, sd in zip(seg, side) if sd == 1]
        perm[s:s + c] = left + right
        if left:
            bucket(s, len(left), lv + 1, r0, c0)
        if right:
            bucket(s + len(left), len(right), lv + 1, r0, c0)

    bucket(0, n, 0, 0, n)
    return np.array(perm, np.int32), np.array(rs, np.int32), np.array(rc, np.int32)


def test_fmaf_restatement():
    """The test's fmaf against exact rational arithmetic, midpoints included."""
    from fractions import Fraction
    rng = np.random.default_rng(0)
    cases = [(np.float32(1 + 2 ** -23), np.float32(1 + 2 ** -23), np.float32(-1.0)),
             (np.float32(3.0), np.float32(2 ** -25), np.float32(1.0)),
             (np.float32(2 ** -24 + 2 ** -47), np.float32(1.0), np.float32(1.0))]
    cases += [tuple(np.float32(v) for v in rng.normal(size=3) * 10.0 ** rng.integers(-3, 4, 3)) for _ in range(300)]
    for a, b, c in cases:
        exact = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
        got = fmaf(a, b, c)
        lo = np.nextafter(got, np.float32(-np.inf))
        hi = np.nextafter(got, np.float32(np.inf))
        err = abs(Fraction(float(got)) - exact)
        assert err <= abs(Fraction(float(lo)) - exact) and err <= abs(Fraction(float(hi)) - exact)


def test_directions_match_restatement(oracle):
    for seed, tree, count, d in ((1, 0, 3, 5), (7, 4, 2, 9), (123456, 11, 1, 1)):
        got = oracle.directions(seed, tree, count, d)
        want = py_directions(seed, tree, count, d)
        assert (got.view(np.uint32) == want.view(np.uint32)).all()
    # a different tree or seed gives different directions
    assert not np.array_equal(oracle.directions(1, 0, 2, 8), oracle.directions(1, 1, 2, 8))
    assert not np.array_equal(oracle.directions(1, 0, 2, 8), oracle.directions(2, 0, 2, 8))
    # roughly standard normal
    r = oracle.directions(3, 0, 200, 100)
    assert abs(r.mean()) < 0.02 and abs(r.std() - 1.0) < 0.02


@pytest.mark.parametrize("n,d,leaf,k", [(700, 3, 20, 5), (1500, 7, 9, 3), (300, 2, 1, 1), (2000, 5, 600, 10)])
def test_median_tree_matches_restatement(oracle, n, d, leaf, k):
    rng = np.random.default_rng(n + d)
    x = rng.normal(size=(n, d)).astype(np.float32)
    x[::7] = x[3]  # duplicates: ties in the projections
    dirs = oracle.directions(5, 0, oracle.median_depth(n, leaf), d)
    perm, rs, rc, lv = oracle.build_tree_median(x, leaf, k, dirs)
    pp, prs, prc = py_median_tree(x, leaf, k, dirs)
    assert (perm == pp).all() and (rs == prs).all() and (rc == prc).all()


def test_routed_tree_matches_restatement(oracle):
    rng = np.random.default_rng(9)
    n, d = 900, 4
    x = rng.normal(size=(n, d)).astype(np.float32)
    bins = rng.integers(0, 8, n).astype(np.int32)
    bins[bins == 5] = 4  # an empty bin
    dirs = oracle.directions(2, 1, oracle.median_depth(n, 16), d)
    perm, rs, rc, lv = oracle.build_tree_median(x, 16, 4, dirs, bins=bins, bin_bits=3)
    pp, prs, prc = py_median_tree(x, 16, 4, dirs, bins=bins, bin_bits=3)
    assert (perm == pp).all() and (rs == prs).all() and (rc == prc).all()
    # bins stay apart: perm is sorted by bin, and no start range of a point
    # with at least k+1 bin-mates mixes bins
    assert (np.diff(bins[perm]) >= 0).all()
    for i in range(n):
        if (bins == bins[i]).sum() >= 5:
            assert (bins[perm[rs[i]:rs[i] + rc[i]]] == bins[i]).all()


@pytest.mark.parametrize("n", [300, 511, 1000, 1021])
def test_exact_halving_when_all_sampled(oracle, n):
    """Buckets of c <= 511 points split into floor(c/2) and ceil(c/2)."""
    rng = np.random.default_rng(n)
    x = rng.normal(size=(n, 6)).astype(np.float32)
    dirs = oracle.directions(1, 0, oracle.median_depth(n, 1), 6)
    leaf = 40
    perm, rs, rc, lv = oracle.build_tree_median(x, leaf, 3, dirs)
    assert (np.sort(perm) == np.arange(n)).all()
    # leaf sizes: repeated halving of n (and of the sampled top split when n > 511)
    sizes = set()

    def halve(c):
        if c <= leaf:
            sizes.add(c)
            return
        halve(c // 2)
        halve(c - c // 2)

    if n <= 511:
        halve(n)
        assert set(np.unique(rc)) <= sizes
    assert rc.max() <= leaf


def test_identical_points_still_split(oracle):
    x = np.ones((200, 3), np.float32)
    dirs = oracle.directions(1, 0, oracle.median_depth(200, 10), 3)
    perm, rs, rc, lv = oracle.build_tree_median(x, 10, 2, dirs)
    assert (np.sort(perm) == np.arange(200)).all() and rc.max() <= 10 and rc.min() >= 3


def test_median_forest_recall_close_to_sign_forest(oracle, datasets):
    """A balanced forest of equal size finds most true neighbours on clustered data."""
    x = datasets.generate("C5", n=8000)
    q = np.arange(0, 8000, 16)
    rec = {}
    for split in (0, 1):
        xn, ids, vals, _ = oracle.run_ex(x, 10, 6, 32, 1, oracle.EUCLIDEAN, split=split)
        rec[split] = oracle.recall(xn, 10, oracle.EUCLIDEAN, q, oracle.exact_kth(xn, 10, oracle.EUCLIDEAN, q), ids)
    assert rec[1] > 0.85, rec


# ---- the HIP build against the oracle ----------------------------------------

def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _gpu_vs_oracle(aknn, oracle, x, k, trees, leaf, props, metric, bins=None, bin_bits=0, router_trees=0,
                   **opts):
    print(f"[gpu] median n={len(x)} d={x.shape[1]} k={k} T={trees} L={leaf} P={props} metric={metric} "
          f"routed={router_trees} {opts}", flush=True)
    g = aknn.Graph()
    g.initialize(x, k, trees, leaf, props, metric, split=1, router_trees=router_trees, **opts)
    if bins is not None:
        g.set_bins(bins, bin_bits)
    g.calculate()
    cf = opts.get("cosine_fixed", 0)
    xn, ids, vals, _ = oracle.run_ex(x, k, trees, leaf, props, metric, cosine_fixed=cf, split=1, bins=bins,
                                     bin_bits=bin_bits, router_trees=router_trees)
    assert (_bits(g.normalized()) == _bits(xn)).all(), "normalized points"
    perm, rs, rc, _ = oracle.build_forest_median(xn, trees, leaf, k, bins=bins, bin_bits=bin_bits,
                                                 router_trees=router_trees)
    for t in range(trees):
        gp, grs, grc, _ = g.forest(t)
        assert (gp == perm[t]).all(), f"tree {t} leaf order"
        assert (grs == rs[t]).all() and (grc == rc[t]).all(), f"tree {t} start ranges"
    gi, gv = g.graph()
    bad = np.nonzero((gi != ids).any(1) | (_bits(gv) != _bits(vals)).any(1))[0]
    assert len(bad) == 0, f"{len(bad)} rows differ, first {bad[:5]}"
    return g


@pytest.mark.gpu
@pytest.mark.parametrize("metric", [0, 1])
def test_gpu_median_forest_bit_exact(aknn, oracle, datasets, metric):
    x = datasets.gmm_c1(3000, 24, clusters=6, seed=3)[0]
    _gpu_vs_oracle(aknn, oracle, x, 8, 4, 20, 2, metric)


@pytest.mark.gpu
@pytest.mark.parametrize("d", [3, 130, 300])
def test_gpu_median_forest_widths(aknn, oracle, datasets, d):
    """Row widths that are not multiples of 4, and D > 256 (two LDS chunks)."""
    x = datasets.sparse_c3(1500, d, clusters=6, seed=d) if d > 3 else datasets.uniform_c2(1500, 3, seed=8)
    _gpu_vs_oracle(aknn, oracle, x, 6, 3, 30, 1, 0)


@pytest.mark.gpu
@pytest.mark.parametrize("leaf,k", [(0, 3), (1, 2), (600, 12), (5000, 9)])
def test_gpu_median_forest_leaf_edges(aknn, oracle, datasets, leaf, k):
    x = datasets.uniform_c2(1500, 12, seed=leaf + 2)
    x[100:400] = x[50]  # a block of identical points
    _gpu_vs_oracle(aknn, oracle, x, k, 2, leaf, 1, 0)


@pytest.mark.gpu
def test_gpu_routed_trees_bit_exact(aknn, oracle, datasets):
    x = datasets.generate("C5", n=5000)
    rng = np.random.default_rng(4)
    bins = rng.integers(0, 16, len(x)).astype(np.int32)
    bins[bins == 7] = 6  # an empty bin
    bins[:40] = 15  # bin 15 gets the first 40 points as well
    _gpu_vs_oracle(aknn, oracle, x, 10, 3, 32, 2, 0, bins=bins, bin_bits=4, router_trees=2)


@pytest.mark.gpu
@pytest.mark.parametrize("cosine_fixed", [0, 1])
def test_gpu_median_forest_cosine(aknn, oracle, datasets, cosine_fixed):
    x = datasets.gmm_c1(2000, 20, clusters=5, seed=4)[0]
    _gpu_vs_oracle(aknn, oracle, x, 6, 3, 24, 1, 2, cosine_fixed=cosine_fixed)


@pytest.mark.gpu
def test_gpu_median_forest_screened_forms_identical(aknn, datasets):
    """The sketch and matrix-core screens leave median-forest graphs unchanged."""
    x = datasets.generate("C5", n=30000)
    g = aknn.Graph()
    g.initialize(x, 20, 4, 64, 2, 0, split=1)
    out = []
    for opts in ({"sketch": 0, "leaf_mfma": 0}, {"sketch": 1, "leaf_mfma": 1}):
        for kk, v in opts.items():
            g.set_option(kk, v)
        print(f"[gpu] median forms {opts}", flush=True)
        g.calculate()
        out.append(g.graph())
    assert (out[0][0] == out[1][0]).all() and (_bits(out[0][1]) == _bits(out[1][1])).all()
