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
    "aknn_get_kernel_stats", "aknn_get_candidate_counts", "aknn_exact_knn", "aknn_recall", "aknn_last_error",
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
    _i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
    L.aknn_exact_knn.argtypes = [H, _i64p, ctypes.c_int64, ctypes.c_int, _i32p, _f64p]
    L.aknn_recall.argtypes = [H, _i64p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]
    L.aknn_last_error.argtypes = [H]
    L.aknn_last_error.restype = ctypes.c_char_p
    for name in EXPORTS:
        if name not in ("aknn_create", "aknn_destroy", "aknn_last_error"):
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

    def deinitialize(self):
        self._check(self._L.aknn_deinitialize(self._h))


def build_graph(x, k, trees, leaf, props, metric, device=0, **options):
    """One-shot helper: returns (ids, vals, stage_times)."""
    with Graph() as g:
        g.initialize(x, k, trees, leaf, props, metric, device=device, **options)
        g.calculate()
        ids, vals = g.graph()
        return ids, vals, g.stage_times()
