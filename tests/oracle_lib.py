"""ctypes binding of the CPU oracle (oracle/liboracle.so).

The oracle is the checker: it is loaded only by the tests, the smoke check
and bench.py's CPU-baseline leg, never by the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

EUCLIDEAN, TAXICAB, COSINE = 0, 1, 2
TAXICAB_L1 = 3  # fixed mode: the L1 total without the square root
METRICS = {"EUCLIDEAN": EUCLIDEAN, "TAXICAB": TAXICAB, "COSINE": COSINE}

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class Rng(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32 * 31), ("f", ctypes.c_int), ("r", ctypes.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_rng_seed.argtypes = [ctypes.POINTER(Rng), ctypes.c_uint32]
        L.orc_rng_next.argtypes = [ctypes.POINTER(Rng)]
        L.orc_rng_next.restype = ctypes.c_int32
        L.orc_pick.argtypes = [ctypes.c_int32, ctypes.c_int32]
        L.orc_pick.restype = ctypes.c_int32
        L.orc_normalize.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int]
        L.orc_distance.argtypes = [_f32p, _f32p, ctypes.c_int, ctypes.c_int]
        L.orc_distance.restype = ctypes.c_float
        L.orc_build_tree.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(Rng), ctypes.c_int, _i32p, _i32p, _i32p,
                                     ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_int64)]
        L.orc_initial_knn.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, _i32p, _i32p, _i32p, _i32p, _f32p, ctypes.c_int]
        L.orc_propagate.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    _i32p, _f32p, _i32p, _f32p, ctypes.c_int]
        L.orc_run.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                              ctypes.c_int, _i32p, _f32p, _f64p]
        L.orc_directions.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p]
        L.orc_median_depth.argtypes = [ctypes.c_int64, ctypes.c_int]
        L.orc_build_tree_median.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            _f32p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, _i32p, _i32p,
                                            _i32p, ctypes.POINTER(ctypes.c_int32)]
        L.orc_run_ex.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, _i32p, _f32p, _f64p]
        L.orc_normalize_ex.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        L.orc_normalize_mode.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_mlp.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              _f32p, _f32p, ctypes.c_float, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                              ctypes.c_void_p]
        L.orc_mlp.restype = ctypes.c_double
        L.orc_exact_kth.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    _i64p, ctypes.c_int64, _f64p]
        L.orc_distance64.argtypes = [_f32p, _f32p, ctypes.c_int, ctypes.c_int]
        L.orc_distance64.restype = ctypes.c_double
        L.orc_recall.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 _i64p, ctypes.c_int64, _f64p, _i32p]
        L.orc_recall.restype = ctypes.c_double
        L.orc_write_result.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                       _i32p, _f32p]
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def rng(seed=1):
    g = Rng()
    lib().orc_rng_seed(ctypes.byref(g), seed)
    return g


def normalize(x, scale_all=0, mean_fp64=0):
    x = np.ascontiguousarray(x, dtype=np.float32).copy()
    lib().orc_normalize_mode(x, x.shape[0], x.shape[1], scale_all, mean_fp64)
    return x


def distance(a, b, metric):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return np.float32(lib().orc_distance(a, b, a.shape[0], metric))


def build_forest(x, trees, leaf, k, seed=1, rng_order=0, g=None):
    """Returns perm, rs, rc as [trees][n] int32 arrays and per-tree iteration counts."""
    n, d = x.shape
    if g is None:
        g = rng(seed)
    perm = np.zeros((trees, n), np.int32)
    rs = np.zeros((trees, n), np.int32)
    rc = np.zeros((trees, n), np.int32)
    iters = []
    for t in range(trees):
        it = ctypes.c_int32()
        nn = ctypes.c_int64()
        act = ctypes.c_int64()
        r = lib().orc_build_tree(x, n, d, leaf, k, ctypes.byref(g), rng_order, perm[t], rs[t], rc[t],
                                 ctypes.byref(it), ctypes.byref(nn), ctypes.byref(act))
        if r != 0:
            raise ValueError("orc_build_tree failed")
        iters.append((it.value, nn.value, act.value))
    return perm, rs, rc, iters


def directions(seed, tree, count, d):
    r = np.zeros((count, d), np.float32)
    lib().orc_directions(seed, tree, count, d, r)
    return r


def median_depth(n, leaf):
    return int(lib().orc_median_depth(n, leaf))


def _bins_ptr(bins):
    if bins is None:
        return None, None
    b = np.ascontiguousarray(bins, dtype=np.int32)
    return b, b.ctypes.data_as(ctypes.c_void_p)


def build_tree_median(x, leaf, k, dirs, bins=None, bin_bits=0):
    """One median-split tree: (perm, rs, rc, levels)."""
    n, d = x.shape
    perm = np.zeros(n, np.int32)
    rs = np.zeros(n, np.int32)
    rc = np.zeros(n, np.int32)
    lv = ctypes.c_int32()
    dirs = np.ascontiguousarray(dirs, dtype=np.float32)
    keep, bp = _bins_ptr(bins)
    r = lib().orc_build_tree_median(x, n, d, leaf, k, dirs, dirs.shape[0], bp, bin_bits, perm, rs, rc,
                                    ctypes.byref(lv))
    if r != 0:
        raise ValueError("orc_build_tree_median failed")
    return perm, rs, rc, lv.value


def build_forest_median(x, trees, leaf, k, seed=1, bins=None, bin_bits=0, router_trees=0):
    """Median-split forest: perm, rs, rc as [trees][n] and per-tree level counts."""
    n, d = x.shape
    nd = median_depth(n, leaf)
    perm = np.zeros((trees, n), np.int32)
    rs = np.zeros((trees, n), np.int32)
    rc = np.zeros((trees, n), np.int32)
    levels = []
    for t in range(trees):
        p, a, c, lv = build_tree_median(x, leaf, k, directions(seed, t, nd, d),
                                        bins if t < router_trees else None, bin_bits)
        perm[t], rs[t], rc[t] = p, a, c
        levels.append(lv)
    return perm, rs, rc, levels


def initial_knn(x, k, metric, perm, rs, rc, cosine_fixed=0):
    n, d = x.shape
    ids = np.zeros((n, k), np.int32)
    vals = np.zeros((n, k), np.float32)
    r = lib().orc_initial_knn(x, n, d, k, perm.shape[0], metric, np.ascontiguousarray(perm),
                              np.ascontiguousarray(rs), np.ascontiguousarray(rc), ids, vals, cosine_fixed)
    if r != 0:
        raise ValueError("orc_initial_knn failed")
    return ids, vals


def propagate(x, ids, vals, metric, cosine_fixed=0):
    n, d = x.shape
    k = ids.shape[1]
    io = np.zeros_like(ids)
    vo = np.zeros_like(vals)
    r = lib().orc_propagate(x, n, d, k, metric, np.ascontiguousarray(ids), np.ascontiguousarray(vals),
                            io, vo, cosine_fixed)
    if r != 0:
        raise ValueError("orc_propagate failed")
    return io, vo


def run(x_raw, k, trees, leaf, props, metric, seed=1, rng_order=0, cosine_fixed=0):
    """Whole pipeline.  Returns (normalized x, ids, vals, stage_ms)."""
    x = np.ascontiguousarray(x_raw, dtype=np.float32).copy()
    n, d = x.shape
    ids = np.zeros((n, k), np.int32)
    vals = np.zeros((n, k), np.float32)
    ms = np.zeros(4, np.float64)
    r = lib().orc_run(x, n, d, k, trees, leaf, props, metric, seed, rng_order, cosine_fixed, ids, vals, ms)
    if r != 0:
        raise ValueError("orc_run failed")
    return x, ids, vals, ms


def run_ex(x_raw, k, trees, leaf, props, metric, seed=1, rng_order=0, cosine_fixed=0, split=0, bins=None,
           bin_bits=0, router_trees=0, scale_all=0):
    """Whole pipeline with the forest rule (split 0 reference, 1 median)."""
    x = np.ascontiguousarray(x_raw, dtype=np.float32).copy()
    n, d = x.shape
    ids = np.zeros((n, k), np.int32)
    vals = np.zeros((n, k), np.float32)
    ms = np.zeros(4, np.float64)
    keep, bp = _bins_ptr(bins)
    r = lib().orc_run_ex(x, n, d, k, trees, leaf, props, metric, seed, rng_order, cosine_fixed, split, bp,
                         bin_bits, router_trees, scale_all, ids, vals, ms)
    if r != 0:
        raise ValueError("orc_run_ex failed")
    return x, ids, vals, ms


def mlp(x, dims, params, mu, inv_sigma, labels=None, round_bf16=True, want_grads=True):
    """Router MLP in FP64: (loss, grads or None, logits).  dims = (d, dp, hidden, bins)."""
    d, dp, h, nb = dims
    x = np.ascontiguousarray(x, dtype=np.float32)
    m = len(x)
    params = np.ascontiguousarray(params, dtype=np.float32)
    mu = np.ascontiguousarray(mu, dtype=np.float32)
    z = np.zeros((m, nb), np.float32)
    g = np.zeros(len(params), np.float32) if (labels is not None and want_grads) else None
    lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.int32)
    loss = lib().orc_mlp(x, m, d, dp, h, nb, params, mu, inv_sigma,
                         None if lab is None else lab.ctypes.data_as(ctypes.c_void_p), 1 if round_bf16 else 0,
                         None if g is None else g.ctypes.data_as(ctypes.c_void_p), z.ctypes.data_as(ctypes.c_void_p))
    return loss, g, z


def exact_kth(x, k, metric, queries):
    n, d = x.shape
    q = np.ascontiguousarray(queries, dtype=np.int64)
    out = np.zeros(len(q), np.float64)
    lib().orc_exact_kth(x, n, d, k, metric, q, len(q), out)
    return out


def recall(x, k, metric, queries, kth, ids):
    n, d = x.shape
    q = np.ascontiguousarray(queries, dtype=np.int64)
    return lib().orc_recall(x, n, d, k, metric, q, len(q), np.ascontiguousarray(kth),
                            np.ascontiguousarray(ids, dtype=np.int32))


def write_result(path, ids, vals, metric):
    n, k = ids.shape
    r = lib().orc_write_result(path.encode(), n, k, metric, np.ascontiguousarray(ids, dtype=np.int32),
                               np.ascontiguousarray(vals, dtype=np.float32))
    if r != 0:
        raise IOError("orc_write_result failed")
