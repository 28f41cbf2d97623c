"""The reference's sign-split forest: the oracle's tree pinned by an
independent restatement.

orc_build_tree restates buildUpTheTrees (euclideandistancematrixgpu.cu:
2407-2686) with per-position arrays, as the reference's kernels hold them.
py_sign_tree below restates the same text a second way, with an explicit node
table {parent, left, right, count} (TreeNode, euclideandistancematrixgpu.h:
28-33) and a bucket list, and with glibc's own rand() (ctypes) as the stream:

  * Leaf1 (.cu:2138-2164): only buckets with size > minSize compute sides;
    side = LEFT when the FP32 chain x . (x_p2 - x_p1) is < 0, else RIGHT;
  * Leaf2 (.cu:2166-2181): side[p1] = LEFT, then side[p2] = RIGHT;
  * Leaf3-6 (.cu:2183-2289): LEFT elements then RIGHT ones (the stable order
    is the legal execution the oracle fixes for the atomic scatter);
  * Leaf7 (.cu:2291-2365, :2589): one (rand(), rand()) pair per iteration,
    the last non-splitting one included; each child of c > 1 points takes
    q1 = r1 % c, q2 = r2 % c (q2 = (q1+1) % c when equal), c == 1 takes (0, 0);
  * Leaf9 (.cu:2391-2405): a leaf node's left = its first position, right = -1;
  * findGraphTraversalStartPoint (.cu:87-105): the first ancestor-or-self of
    the point's leaf node with count >= k+1; its range starts at the
    leftmost leaf below it (.cu:166-176).

The two restatements are compared bit for bit on trees at least 4 levels
deep, and the restatement's node table is checked against what the reference
text guarantees: leaves of 1..minSize points, internal nodes of more than
minSize points with children summing to them, p1 ending LEFT and p2 RIGHT at
every split, start nodes minimal.  Three plausible misreadings planted in the
restatement (size >= minSize, a child's pivots drawn with its parent's count,
a start node one level too high) each make the comparison fail.
"""
import ctypes

import numpy as np
import pytest

from test_median_forest import fmaf

LEFT, RIGHT = 0, 2  # .cu:9-11

_libc = ctypes.CDLL("libc.so.6")
_libc.rand.restype = ctypes.c_int
_libc.srand.argtypes = [ctypes.c_uint]


def pick(r, count):
    """int(((double)r / RAND_MAX) * INT_MAX) % count (.cu:2482-2484, :2322-2323)."""
    return int((r / 2147483647.0) * 2147483647.0) % count


def side_chain(xc, xa, xb):
    """sideSign += x_c * (x_p2 - x_p1) over the dimensions in order, one FP32
    FMA per step (.cu:2153-2157)."""
    s = np.float32(0.0)
    for c, a, b in zip(xc, xa, xb):
        s = fmaf(c, np.float32(b - a), s)
    return s


def py_sign_tree(x, leaf_max, k, rng_order=0, slip=None, trace=None):
    """One tree of the reference's forest, drawing from libc rand() (srand
    before the first tree).  Returns perm, range_start, range_count (by point
    id), iterations, node count, active elements and the node table."""
    n = len(x)
    # node table: [parent, left, right, count]; the root is node 0
    nodes = [[-1, -1, -1, n]]
    a, b = _libc.rand(), _libc.rand()
    q1, q2 = pick(a, n), pick(b, n)
    if q1 == q2:
        q2 = (q1 + 1) % n
    perm = list(range(n))
    buckets = [{"start": 0, "size": n, "node": 0, "q1": q1, "q2": q2}]
    iters = 0
    active = 0
    while True:
        iters += 1
        split_any = False
        nxt = []
        splits = []
        for bk in buckets:
            s0, c = bk["start"], bk["size"]
            splits_here = c >= leaf_max if slip == "ge_leaf" else c > leaf_max
            if not splits_here:
                nxt.append(bk)
                continue
            active += c
            p1, p2 = perm[s0 + bk["q1"]], perm[s0 + bk["q2"]]
            side = []
            for pos in range(s0, s0 + c):
                s = side_chain(x[perm[pos]], x[p1], x[p2])
                side.append(LEFT if s < 0 else RIGHT)
            side[bk["q1"]] = LEFT
            side[bk["q2"]] = RIGHT
            seg = perm[s0:s0 + c]
            left = [e for e, sd in zip(seg, side) if sd == LEFT]
            right = [e for e, sd in zip(seg, side) if sd != LEFT]
            perm[s0:s0 + c] = left + right
            if left and right:
                splits.append((bk, len(left), len(right), p1, p2))
            else:
                nxt.append(bk)
        first, second = _libc.rand(), _libc.rand()
        r1, r2 = (second, first) if rng_order == 0 else (first, second)
        for bk, nl, nr, p1, p2 in splits:
            split_any = True
            parent = bk["node"]
            for cnt, start in ((nl, bk["start"]), (nr, bk["start"] + nl)):
                nodes.append([parent, -1, -1, cnt])
                pc = bk["size"] if slip == "parent_count" else cnt
                if cnt > 1:
                    i1, i2 = pick(r1, pc) % cnt, pick(r2, pc) % cnt
                    if i1 == i2:
                        i2 = (i1 + 1) % cnt
                else:
                    i1 = i2 = 0
                nxt.append({"start": start, "size": cnt, "node": len(nodes) - 1, "q1": i1, "q2": i2})
            nodes[parent][1], nodes[parent][2] = len(nodes) - 2, len(nodes) - 1
            if trace is not None:
                trace.append((parent, p1, p2, nl, nr))
        buckets = sorted(nxt, key=lambda q: q["start"])
        if not split_any:
            break
    leaf_of = {}
    for bk in buckets:
        nodes[bk["node"]][1], nodes[bk["node"]][2] = bk["start"], -1
        for pos in range(bk["start"], bk["start"] + bk["size"]):
            leaf_of[perm[pos]] = bk["node"]
    rs = np.zeros(n, np.int32)
    rc = np.zeros(n, np.int32)
    for pid in range(n):
        v = leaf_of[pid]
        if slip == "start_high":
            while nodes[v][3] <= k + 1:
                v = nodes[v][0]
        else:
            while nodes[v][3] < k + 1:
                v = nodes[v][0]
        u = v
        while nodes[u][2] != -1:
            u = nodes[u][1]
        rs[pid], rc[pid] = nodes[u][1], nodes[v][3]
    return np.array(perm, np.int32), rs, rc, iters, len(nodes), active, nodes, leaf_of


def py_sign_forest(x, trees, leaf_max, k, seed=1, rng_order=0, slip=None, traces=None):
    _libc.srand(seed)
    out = []
    for _ in range(trees):
        tr = [] if traces is not None else None
        out.append(py_sign_tree(x, leaf_max, k, rng_order, slip, tr))
        if traces is not None:
            traces.append(tr)
    return out


# small inputs whose trees are at least 4 levels deep: (n, d, leaf, k, seed,
# rng_order, trees, data kind)
CASES = [
    (60, 1, 1, 2, 1, 0, 2, "normal"),
    (90, 2, 3, 4, 7, 1, 2, "normal"),
    (120, 3, 7, 5, 3, 0, 2, "clusters"),
    (75, 2, 1, 1, 11, 1, 3, "clusters"),
    (150, 3, 3, 9, 5, 0, 2, "offset"),
    (110, 1, 7, 3, 2, 1, 2, "ints"),
    (64, 3, 1, 6, 9, 0, 2, "dups"),
    (130, 2, 7, 0, 4, 1, 2, "normal"),
]


def _data(oracle, n, d, kind, seed):
    rng = np.random.default_rng(100 + seed)
    if kind == "normal":
        x = rng.normal(size=(n, d))
    elif kind == "clusters":
        c = rng.normal(0, 5, (4, d))
        x = c[rng.integers(0, 4, n)] + rng.normal(0, 0.5, (n, d))
    elif kind == "offset":  # far from the origin: sign splits peel points off
        x = rng.normal(size=(n, d)) + 3.0
    elif kind == "ints":
        x = rng.integers(-4, 5, (n, d)).astype(np.float64)
    else:  # duplicated rows: exact zeros in the side chain
        x = np.repeat(rng.normal(size=(n // 4, d)), 4, axis=0)
    return oracle.normalize(x.astype(np.float32))


def _oracle_forest(oracle, x, trees, leaf, k, seed, order):
    return oracle.build_forest(x, trees, leaf, k, seed=seed, rng_order=order)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"n{c[0]}d{c[1]}L{c[2]}k{c[3]}s{c[4]}o{c[5]}{c[7]}")
def test_restatement_matches_oracle(oracle, case):
    n, d, leaf, k, seed, order, trees, kind = case
    x = _data(oracle, n, d, kind, seed)
    perm, rs, rc, info = _oracle_forest(oracle, x, trees, leaf, k, seed, order)
    model = py_sign_forest(x, trees, leaf, k, seed, order)
    for t in range(trees):
        p, a, c, it, nn, act = model[t][:6]
        assert it >= 5, "the case must build a tree at least 4 levels deep"
        assert (perm[t] == p).all(), f"tree {t}: leaf order differs"
        assert (rs[t] == a).all() and (rc[t] == c).all(), f"tree {t}: start ranges differ"
        assert info[t] == (it, nn, act), f"tree {t}: iterations / nodes / active elements differ"


@pytest.mark.parametrize("case", CASES[:5], ids=lambda c: f"n{c[0]}d{c[1]}L{c[2]}k{c[3]}")
def test_restatement_invariants(oracle, case):
    n, d, leaf, k, seed, order, trees, kind = case
    x = _data(oracle, n, d, kind, seed)
    traces = []
    model = py_sign_forest(x, trees, leaf, k, seed, order, traces=traces)
    for t in range(trees):
        perm, rs, rc, it, nn, act, nodes, leaf_of = model[t]
        assert sorted(perm.tolist()) == list(range(n))
        children = {}
        for v, (par, l, r, cnt) in enumerate(nodes):
            if par >= 0:
                children.setdefault(par, []).append(v)
        for v, (par, l, r, cnt) in enumerate(nodes):
            if v in children:
                # internal: more than minSize points, two children summing to it
                assert cnt > leaf and len(children[v]) == 2
                assert sum(nodes[c][3] for c in children[v]) == cnt
                assert r != -1
            else:
                assert 1 <= cnt <= max(leaf, 1) and r == -1
        # p1 ends in the left child and p2 in the right one at every split
        pos = {int(pid): i for i, pid in enumerate(perm)}
        for parent, p1, p2, nl, nr in traces[t]:
            lc, rcn = nodes[parent][1], nodes[parent][2]
            assert nodes[lc][3] == nl and nodes[rcn][3] == nr

            def under(v, pid):
                u = leaf_of[pid]
                while u != -1 and u != v:
                    u = nodes[u][0]
                return u == v

            assert under(lc, p1) and under(rcn, p2)
        # start node: count >= k+1, and the child on the point's path holds
        # fewer than k+1 points (the first such ancestor)
        for pid in range(n):
            u = leaf_of[pid]
            path = [u]
            while nodes[u][0] != -1:
                u = nodes[u][0]
                path.append(u)
            first = next(v for v in path if nodes[v][3] >= k + 1)
            assert rc[pid] == nodes[first][3]
            assert rs[pid] <= pos[pid] < rs[pid] + rc[pid]


def test_leaf_sizes_from_oracle(oracle):
    """With k = 0 a start node is the point's own leaf, so the oracle's start
    ranges are its leaves: every leaf holds 1..minSize points, the leaves tile
    the permutation, and the bucket above (k = 1 on one-point leaves) is larger
    than minSize."""
    for n, d, leaf, seed in ((300, 3, 1, 1), (400, 2, 3, 2), (500, 3, 7, 3), (200, 1, 5, 4)):
        x = _data(oracle, n, d, "normal", seed)
        perm, rs, rc, _ = _oracle_forest(oracle, x, 2, leaf, 0, seed, 0)
        for t in range(2):
            assert ((rc[t] >= 1) & (rc[t] <= leaf)).all()
            starts = sorted(set(zip(rs[t].tolist(), rc[t].tolist())))
            cur = 0
            for s, c in starts:
                assert s == cur
                cur += c
            assert cur == n


@pytest.mark.parametrize("slip", ["ge_leaf", "parent_count", "start_high"])
def test_planted_slips_are_caught(oracle, slip):
    """Each misreading, planted in the restatement, disagrees with the oracle
    on at least one of the cases."""
    caught = False
    for n, d, leaf, k, seed, order, trees, kind in CASES:
        x = _data(oracle, n, d, kind, seed)
        perm, rs, rc, info = _oracle_forest(oracle, x, trees, leaf, k, seed, order)
        model = py_sign_forest(x, trees, leaf, k, seed, order, slip=slip)
        for t in range(trees):
            p, a, c, it, nn, act = model[t][:6]
            if not ((perm[t] == p).all() and (rs[t] == a).all() and (rc[t] == c).all() and
                    info[t] == (it, nn, act)):
                caught = True
        if caught:
            break
    assert caught


def test_libc_rand_is_the_oracle_stream(oracle):
    g = oracle.rng(5)
    _libc.srand(5)
    for _ in range(1000):
        assert oracle.lib().orc_rng_next(ctypes.byref(g)) == _libc.rand()
