"""ctypes binding of libaknn.so (the C ABI in include/aknn.h).

This is the Python face of the drop-in boundary: the same lifecycle as the
reference's EuclideanDistanceMatrixGPU (euclideandistancematrixgpu.h:50-63).
It loads the in-tree build and raises if it is missing; there is no fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaknn.so")
CLI_PATH = os.path.join(HERE, "aknn")

EUCLIDEAN, TAXICAB, COSINE = 0, 1, 2
METRICS = {"EUCLIDEAN": EUCLIDEAN, "TAXICAB": TAXICAB, "COSINE": COSINE}

# Every symbol include/aknn.h declares.
EXPORTS = (
    "aknn_create", "aknn_destroy", "aknn_set_data_file", "aknn_set_results_file", "aknn_initialize",
    "aknn_initialize_host", "aknn_calculate", "aknn_save_result", "aknn_deinitialize", "aknn_set_option",
    "aknn_get_graph", "aknn_get_normalized", "aknn_get_forest", "aknn_get_stage_times",
    "aknn_get_kernel_stats", "aknn_get_candidate_counts", "aknn_exact_knn", "aknn_recall", "aknn_set_bins",
    "aknn_last_error", "aknn_router_create", "aknn_router_destroy", "aknn_router_labels", "aknn_router_fit",
    "aknn_router_gradients", "aknn_router_logits", "aknn_router_param_count", "aknn_router_get_params",
    "aknn_router_set_params", "aknn_router_save", "aknn_router_load", "aknn_router_stats",
    "aknn_router_last_error", "aknn_set_router", "aknn_read_points", "aknn_router_bins", "aknn_add_router",
    "aknn_recall_noalloc", "aknn_get_leaf_form",
)

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")

_lib = None


class AknnError(RuntimeError):
    pass


def load():
    """Loads libaknn.so; raises AknnError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AknnError(f"{LIB_PATH} is missing: build it with `make -C {HERE}` (or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    H = ctypes.c_void_p
    L.aknn_create.argtypes = [ctypes.c_int]
    L.aknn_create.restype = H
    L.aknn_destroy.argtypes = [H]
    L.aknn_destroy.restype = None
    L.aknn_set_data_file.argtypes = [H, ctypes.c_char_p]
    L.aknn_set_results_file.argtypes = [H, ctypes.c_char_p]
    L.aknn_initialize.argtypes = [H] + [ctypes.c_int] * 9
    L.aknn_initialize_host.argtypes = [H, _f32p, ctypes.c_int64] + [ctypes.c_int] * 7
    L.aknn_calculate.argtypes = [H]
    L.aknn_save_result.argtypes = [H]
    L.aknn_deinitialize.argtypes = [H]
    L.aknn_set_option.argtypes = [H, ctypes.c_char_p, ctypes.c_int64]
    L.aknn_get_graph.argtypes = [H, _i32p, _f32p]
    L.aknn_get_normalized.argtypes = [H, _f32p]
    L.aknn_get_forest.argtypes = [H, ctypes.c_int, _i32p, _i32p, _i32p, ctypes.POINTER(ctypes.c_int32)]
    L.aknn_get_stage_times.argtypes = [H, _f64p, ctypes.c_int]
    L.aknn_get_kernel_stats.argtypes = [H, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.aknn_get_candidate_counts.argtypes = [H, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double)]
    L.aknn_get_leaf_form.argtypes = [H, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double)]
    _i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
    L.aknn_exact_knn.argtypes = [H, _i64p, ctypes.c_int64, ctypes.c_int, _i32p, _f64p]
    L.aknn_recall.argtypes = [H, _i64p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]
    L.aknn_recall_noalloc.argtypes = [H, _i64p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]
    L.aknn_set_bins.argtypes = [H, _i32p, ctypes.c_int64, ctypes.c_int]
    L.aknn_last_error.argtypes = [H]
    L.aknn_last_error.restype = ctypes.c_char_p
    R = ctypes.c_void_p
    L.aknn_router_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32]
    L.aknn_router_create.restype = R
    L.aknn_router_destroy.argtypes = [R]
    L.aknn_router_destroy.restype = None
    L.aknn_router_labels.argtypes = [R, H, ctypes.c_int, _i32p, _f64p]
    L.aknn_router_fit.argtypes = [R, H, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float]
    L.aknn_router_gradients.argtypes = [R, _f32p, ctypes.c_int64, _i32p, _f32p, ctypes.POINTER(ctypes.c_double)]
    L.aknn_router_logits.argtypes = [R, _f32p, ctypes.c_int64, _f32p]
    L.aknn_router_bins.argtypes = [R, _f32p, ctypes.c_int64, ctypes.c_int, _i32p]
    L.aknn_router_param_count.argtypes = [R, ctypes.POINTER(ctypes.c_int64)]
    L.aknn_router_get_params.argtypes = [R, _f32p, _f32p, ctypes.POINTER(ctypes.c_float)]
    L.aknn_router_set_params.argtypes = [R, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float]
    L.aknn_router_save.argtypes = [R, ctypes.c_char_p]
    L.aknn_router_load.argtypes = [R, ctypes.c_char_p]
    L.aknn_router_stats.argtypes = [R, _f64p, ctypes.c_int]
    L.aknn_router_last_error.argtypes = [R]
    L.aknn_router_last_error.restype = ctypes.c_char_p
    L.aknn_set_router.argtypes = [H, R]
    L.aknn_add_router.argtypes = [H, R]
    L.aknn_read_points.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   _f32p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
    for name in EXPORTS:
        if name not in ("aknn_create", "aknn_destroy", "aknn_last_error", "aknn_router_create",
                        "aknn_router_destroy", "aknn_router_last_error"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


class Graph:
    """One handle of the library: upload points once, calculate() as often as needed."""

    def __init__(self, debug=False):
        self._L = load()
        self._h = self._L.aknn_create(1 if debug else 0)
        if not self._h:
            raise AknnError("aknn_create failed")
        self.n = self.d = self.k = self.trees = None

    def _check(self, rc):
        if rc != 0:
            raise AknnError(self._L.aknn_last_error(self._h).decode(errors="replace") or f"status {rc}")

    def close(self):
        if self._h:
            self._L.aknn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_option(self, name, value):
        self._check(self._L.aknn_set_option(self._h, name.encode(), int(value)))

    def initialize(self, x, k, trees, leaf, props, metric, device=0, **options):
        x = np.ascontiguousarray(x, dtype=np.float32)
        for name, value in options.items():
            self.set_option(name, value)
        n, d = x.shape
        self._check(self._L.aknn_initialize_host(self._h, x, n, d, k, device, metric, trees, props, leaf))
        self.n, self.d, self.k, self.trees = n, d, k, trees

    def initialize_csv(self, path, n, d, k, trees, leaf, props, metric, klaster=False, device=0):
        self._check(self._L.aknn_set_data_file(self._h, path.encode()))
        self._check(self._L.aknn_initialize(self._h, n, d, k, device, metric, 1 if klaster else 0, trees, props,
                                            leaf))
        self.n, self.d, self.k, self.trees = n, d, k, trees

    def calculate(self):
        self._check(self._L.aknn_calculate(self._h))

    def save(self, path):
        self._check(self._L.aknn_set_results_file(self._h, path.encode()))
        self._check(self._L.aknn_save_result(self._h))

    def graph(self):
        ids = np.empty((self.n, self.k), np.int32)
        vals = np.empty((self.n, self.k), np.float32)
        self._check(self._L.aknn_get_graph(self._h, ids, vals))
        return ids, vals

    def normalized(self):
        x = np.empty((self.n, self.d), np.float32)
        self._check(self._L.aknn_get_normalized(self._h, x))
        return x

    def forest(self, tree):
        perm = np.empty(self.n, np.int32)
        rs = np.empty(self.n, np.int32)
        rc = np.empty(self.n, np.int32)
        it = ctypes.c_int32()
        self._check(self._L.aknn_get_forest(self._h, tree, perm, rs, rc, ctypes.byref(it)))
        return perm, rs, rc, it.value

    def stage_times(self):
        ms = np.zeros(5, np.float64)
        self._check(self._L.aknn_get_stage_times(self._h, ms, 5))
        return dict(zip(("normalize", "forest", "initial", "propagation", "total"), ms.tolist()))

    def kernel_stats(self, name):
        launches = ctypes.c_int64()
        total = ctypes.c_double()
        units = ctypes.c_double()
        self._check(self._L.aknn_get_kernel_stats(self._h, name.encode(), ctypes.byref(launches),
                                                  ctypes.byref(total), ctypes.byref(units)))
        return {"launches": launches.value, "total_ms": total.value, "units": units.value}

    def candidate_counts(self, stage="propagate"):
        """(examined, rescored) candidates of the last calculate(): propagation
        candidates, or for stage "leaf_knn" the pairs screened on the matrix
        cores and those re-scored exactly."""
        ex = ctypes.c_double()
        rs = ctypes.c_double()
        self._check(self._L.aknn_get_candidate_counts(self._h, stage.encode(), ctypes.byref(ex), ctypes.byref(rs)))
        return ex.value, rs.value

    def leaf_form(self):
        """Leaf k-NN launch form of the last calculate() ("all_trees" or
        "per_tree") and the two forms' times (ms) measured by the call that
        chose it (option leaf_batch 2)."""
        f = ctypes.c_int32()
        a = ctypes.c_double()
        b = ctypes.c_double()
        self._check(self._L.aknn_get_leaf_form(self._h, ctypes.byref(f), ctypes.byref(a), ctypes.byref(b)))
        return {"form": "all_trees" if f.value else "per_tree", "tuning_ms": {"per_tree": a.value,
                                                                            "all_trees": b.value}}

    def exact_knn(self, queries, kq):
        """FP64 exact k-NN of the query indices: (ids [nq][kq], dist [nq][kq])."""
        q = np.ascontiguousarray(queries, dtype=np.int64)
        ids = np.empty((len(q), kq), np.int32)
        dist = np.empty((len(q), kq), np.float64)
        self._check(self._L.aknn_exact_knn(self._h, q, len(q), kq, ids, dist))
        return ids, dist

    def recall(self, queries):
        """recall@k of the current graph against FP64 exact k-NN of the queries."""
        q = np.ascontiguousarray(queries, dtype=np.int64)
        r = ctypes.c_double()
        self._check(self._L.aknn_recall(self._h, q, len(q), ctypes.byref(r)))
        return r.value

    def recall_noalloc(self, queries):
        """recall() without device allocation (buffers from the library's
        static device scratch; the same result)."""
        q = np.ascontiguousarray(queries, dtype=np.int64)
        r = ctypes.c_double()
        self._check(self._L.aknn_recall_noalloc(self._h, q, len(q), ctypes.byref(r)))
        return r.value

    def set_bins(self, bins, bin_bits):
        """Router bins [n] of the points (values < 2^bin_bits) for the routed
        trees of the median-split forest."""
        b = np.ascontiguousarray(bins, dtype=np.int32)
        self._check(self._L.aknn_set_bins(self._h, b, len(b), bin_bits))

    def set_router(self, router):
        """Attaches a trained Router (None detaches); routed trees (option
        router_trees, split=1) then take its bins inside calculate()."""
        self._check(self._L.aknn_set_router(self._h, router._r if router is not None else None))
        self._routers = [] if router is None else [router]

    def add_router(self, router):
        """Appends a router: routed tree t takes router t mod R."""
        self._check(self._L.aknn_add_router(self._h, router._r))
        self._routers = getattr(self, "_routers", []) + [router]

    def deinitialize(self):
        self._check(self._L.aknn_deinitialize(self._h))


def read_points(path, n, d, klaster=False, threads=0):
    """The library's input loader without a device: (x [n][d] float32, labels or None)."""
    L = load()
    x = np.empty((n, d), np.float32)
    lab = np.empty(n, np.int32) if klaster else None
    err = ctypes.create_string_buffer(512)
    rc = L.aknn_read_points(path.encode(), n, d, 1 if klaster else 0, threads, x,
                            None if lab is None else lab.ctypes.data_as(ctypes.c_void_p), err, 512)
    if rc != 0:
        raise AknnError(err.value.decode(errors="replace"))
    return x, lab


class Router:
    """The learned router (MLP d -> hidden -> hidden -> 2^bin_bits bins) of the
    median-split forest, trained on the device on a graph this library built."""

    def __init__(self, d, hidden=256, bin_bits=8, seed=1, device=0):
        self._L = load()
        self._r = self._L.aknn_router_create(device, d, hidden, bin_bits, seed)
        if not self._r:
            raise AknnError("aknn_router_create failed (device, D, hidden or bin_bits out of range)")
        self.d, self.hidden, self.bin_bits, self.bins = d, hidden, bin_bits, 1 << bin_bits
        self.dp = (d + 15) // 16 * 16

    def _check(self, rc):
        if rc != 0:
            raise AknnError(self._L.aknn_router_last_error(self._r).decode(errors="replace") or f"status {rc}")

    def close(self):
        if self._r:
            self._L.aknn_router_destroy(self._r)
            self._r = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def labels(self, graph, sweeps=10):
        """Balanced graph labels of graph's points and the (before, after) edge-cut fractions."""
        lab = np.empty(graph.n, np.int32)
        cut = np.zeros(2, np.float64)
        self._check(self._L.aknn_router_labels(self._r, graph._h, sweeps, lab, cut))
        return lab, (float(cut[0]), float(cut[1]))

    def fit(self, graph, labels=None, sweeps=10, epochs=10, batch=1024, lr=2e-3):
        lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.int32)
        ptr = None if lab is None else lab.ctypes.data_as(ctypes.c_void_p)
        self._check(self._L.aknn_router_fit(self._r, graph._h, ptr, sweeps, epochs, batch, lr))
        return self.stats()

    def gradients(self, x, labels):
        x = np.ascontiguousarray(x, dtype=np.float32)
        lab = np.ascontiguousarray(labels, dtype=np.int32)
        g = np.empty(self.param_count(), np.float32)
        loss = ctypes.c_double()
        self._check(self._L.aknn_router_gradients(self._r, x, len(x), lab, g, ctypes.byref(loss)))
        return g, loss.value

    def logits(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        z = np.empty((len(x), self.bins), np.float32)
        self._check(self._L.aknn_router_logits(self._r, x, len(x), z))
        return z

    def bins_of(self, x, fused=True):
        """Bins of raw points through the assignment path calculate() uses."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        b = np.empty(len(x), np.int32)
        self._check(self._L.aknn_router_bins(self._r, x, len(x), 1 if fused else 0, b))
        return b

    def param_count(self):
        c = ctypes.c_int64()
        self._check(self._L.aknn_router_param_count(self._r, ctypes.byref(c)))
        return c.value

    def get_params(self):
        p = np.empty(self.param_count(), np.float32)
        mu = np.empty(self.d, np.float32)
        inv = ctypes.c_float()
        self._check(self._L.aknn_router_get_params(self._r, p, mu, ctypes.byref(inv)))
        return p, mu, inv.value

    def set_params(self, params=None, mu=None, inv_sigma=1.0):
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        m = None if mu is None else np.ascontiguousarray(mu, dtype=np.float32)
        self._check(self._L.aknn_router_set_params(
            self._r, None if p is None else p.ctypes.data_as(ctypes.c_void_p),
            None if m is None else m.ctypes.data_as(ctypes.c_void_p), inv_sigma))

    def split_params(self, p):
        """Views W1 [hidden][dp], b1, W2, b2, W3 [bins][hidden], b3 of a flat parameter array."""
        h, dp, nb = self.hidden, self.dp, self.bins
        sizes = [h * dp, h, h * h, h, nb * h, nb]
        shapes = [(h, dp), (h,), (h, h), (h,), (nb, h), (nb,)]
        out, o = [], 0
        for sz, sh in zip(sizes, shapes):
            out.append(p[o:o + sz].reshape(sh))
            o += sz
        return out

    def save(self, path):
        self._check(self._L.aknn_router_save(self._r, path.encode()))

    def load(self, path):
        self._check(self._L.aknn_router_load(self._r, path.encode()))

    def stats(self):
        v = np.zeros(7, np.float64)
        self._check(self._L.aknn_router_stats(self._r, v, 7))
        keys = ("epochs", "loss", "accuracy", "train_ms", "cut_before", "cut_after", "assign_ms")
        return dict(zip(keys, v.tolist()))


def build_graph(x, k, trees, leaf, props, metric, device=0, **options):
    """One-shot helper: returns (ids, vals, stage_times)."""
    with Graph() as g:
        g.initialize(x, k, trees, leaf, props, metric, device=device, **options)
        g.calculate()
        ids, vals = g.graph()
        return ids, vals, g.stage_times()
