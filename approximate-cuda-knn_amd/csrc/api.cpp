// extern "C" boundary of libaknn.so (declarations and reference citations in
// include/aknn.h).  Every entry point catches everything: nothing throws across
// the ABI.
#include "aknn.h"

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine.hpp"
#include "io.hpp"
#include "router.hpp"

struct aknn_router {
  std::unique_ptr<aknn::Router> router;
  std::string error;
};

struct aknn_handle {
  bool debug = false;
  std::string data_file, results_file, error;
  aknn::Params params;
  bool have_options = false;
  int device = -1;
  std::unique_ptr<aknn::Engine> engine;
  bool calculated = false;
  int io_threads = 0;
};

namespace {

int fail(aknn_handle* h, int code, const std::string& msg) {
  if (h) h->error = msg;
  return code;
}

template <typename F>
int guarded(aknn_handle* h, F&& f) {
  if (!h) return AKNN_ERR_ARGUMENT;
  try {
    h->error.clear();
    return f();
  } catch (const aknn::HipError& e) {
    return fail(h, AKNN_ERR_DEVICE, e.what);
  } catch (const std::invalid_argument& e) {
    return fail(h, AKNN_ERR_ARGUMENT, e.what());
  } catch (const std::bad_alloc&) {
    return fail(h, AKNN_ERR_DEVICE, "out of host memory");
  } catch (const std::exception& e) {
    return fail(h, AKNN_ERR_STATE, e.what());
  } catch (...) {
    return fail(h, AKNN_ERR_STATE, "unknown error");
  }
}

int setup(aknn_handle* h, const float* x, int64_t n, int d, int k, int device, int distance, int trees, int props,
          int min_size) {
  aknn::Params p = h->params;
  p.k = k;
  p.trees = trees;
  p.leaf = min_size;
  p.props = props;
  p.metric = distance;
  if (!h->engine || h->device != device) {
    h->engine.reset();
    h->engine = std::make_unique<aknn::Engine>(device);
    h->device = device;
  }
  h->engine->set_data(x, n, d, p);
  h->params = p;
  h->calculated = false;
  return AKNN_OK;
}

int rfail(aknn_router* r, int code, const std::string& msg) {
  if (r) r->error = msg;
  return code;
}

template <typename F>
int rguarded(aknn_router* r, F&& f) {
  if (!r || !r->router) return AKNN_ERR_ARGUMENT;
  try {
    r->error.clear();
    return f();
  } catch (const aknn::HipError& e) {
    return rfail(r, AKNN_ERR_DEVICE, e.what);
  } catch (const std::invalid_argument& e) {
    return rfail(r, AKNN_ERR_ARGUMENT, e.what());
  } catch (const std::bad_alloc&) {
    return rfail(r, AKNN_ERR_DEVICE, "out of host memory");
  } catch (const std::exception& e) {
    return rfail(r, AKNN_ERR_STATE, e.what());
  } catch (...) {
    return rfail(r, AKNN_ERR_STATE, "unknown error");
  }
}

}  // namespace

extern "C" {

aknn_handle* aknn_create(int debug_mode) {
  try {
    auto* h = new aknn_handle();
    h->debug = debug_mode != 0;
    return h;
  } catch (...) {
    return nullptr;
  }
}

void aknn_destroy(aknn_handle* h) {
  try {
    delete h;
  } catch (...) {
  }
}

int aknn_set_data_file(aknn_handle* h, const char* path) {
  return guarded(h, [&] {
    if (!path) return fail(h, AKNN_ERR_ARGUMENT, "null path");
    h->data_file = path;
    return AKNN_OK;
  });
}

int aknn_set_results_file(aknn_handle* h, const char* path) {
  return guarded(h, [&] {
    if (!path) return fail(h, AKNN_ERR_ARGUMENT, "null path");
    h->results_file = path;
    return AKNN_OK;
  });
}

int aknn_initialize(aknn_handle* h, int n, int d, int k, int device, int distance, int klaster, int trees,
                    int propagations, int min_size) {
  return guarded(h, [&] {
    if (n < 1 || d < 1) return fail(h, AKNN_ERR_ARGUMENT, "N and D must be positive");
    std::vector<float> x;
    std::vector<int32_t> labels;
    // *.fbin: binary points; otherwise the reference's CSV, parsed by
    // io_threads host threads (0: up to 16), the same values as one thread
    const std::string err = aknn::is_fbin(h->data_file)
                                ? aknn::load_fbin(h->data_file, n, d, x)
                                : (h->io_threads == 1 ? aknn::load_csv(h->data_file, n, d, klaster != 0, x, labels)
                                                      : aknn::load_csv_parallel(h->data_file, n, d, klaster != 0,
                                                                                h->io_threads, x, labels));
    if (!err.empty()) return fail(h, AKNN_ERR_IO, err);
    return setup(h, x.data(), n, d, k, device, distance, trees, propagations, min_size);
  });
}

int aknn_initialize_host(aknn_handle* h, const float* x, int64_t n, int d, int k, int device, int distance,
                         int trees, int propagations, int min_size) {
  return guarded(h, [&] {
    if (!x) return fail(h, AKNN_ERR_ARGUMENT, "null data");
    return setup(h, x, n, d, k, device, distance, trees, propagations, min_size);
  });
}

int aknn_calculate(aknn_handle* h) {
  return guarded(h, [&] {
    if (!h->engine) return fail(h, AKNN_ERR_STATE, "not initialized");
    h->engine->calculate();
    h->calculated = true;
    return AKNN_OK;
  });
}

int aknn_save_result(aknn_handle* h) {
  return guarded(h, [&] {
    if (!h->engine || !h->calculated) return fail(h, AKNN_ERR_STATE, "nothing calculated");
    const int64_t nk = h->engine->n() * h->params.k;
    std::vector<int32_t> ids(nk);
    std::vector<float> vals(nk);
    h->engine->get_graph(ids.data(), vals.data());
    const std::string err = aknn::write_result(h->results_file, h->debug, h->params.metric, h->engine->n(),
                                               h->params.k, ids.data(), vals.data());
    if (!err.empty()) return fail(h, AKNN_ERR_IO, err);
    return AKNN_OK;
  });
}

int aknn_deinitialize(aknn_handle* h) {
  return guarded(h, [&] {
    h->engine.reset();
    h->device = -1;
    h->calculated = false;
    return AKNN_OK;
  });
}

int aknn_set_option(aknn_handle* h, const char* name, int64_t value) {
  return guarded(h, [&] {
    if (!name) return fail(h, AKNN_ERR_ARGUMENT, "null option name");
    const std::string n(name);
    if (n == "seed") {
      if (value < 0 || value > 0xffffffffLL) return fail(h, AKNN_ERR_ARGUMENT, "seed must fit in 32 bits");
      h->params.seed = static_cast<uint32_t>(value);
    } else if (n == "rng_order") {
      if (value != 0 && value != 1) return fail(h, AKNN_ERR_ARGUMENT, "rng_order is 0 or 1");
      h->params.rng_order = static_cast<int>(value);
    } else if (n == "cosine_fixed") {
      h->params.cosine_fixed = value != 0;
    } else if (n == "level_batch") {
      if (value < 1 || value > 4096) return fail(h, AKNN_ERR_ARGUMENT, "level_batch must be in [1, 4096]");
      h->params.level_batch = static_cast<int>(value);
    } else if (n == "skip_settled") {
      h->params.skip_settled = value != 0;
    } else if (n == "sketch") {
      h->params.sketch = value != 0;
    } else if (n == "leaf_mfma") {
      h->params.leaf_mfma = value != 0;
    } else if (n == "sketch_int8") {
      h->params.sketch_int8 = value != 0;
    } else if (n == "forest_int8") {
      h->params.forest_int8 = value != 0;
    } else if (n == "flip_int8") {
      h->params.flip_int8 = value != 0;
    } else if (n == "split") {
      if (value != 0 && value != 1) return fail(h, AKNN_ERR_ARGUMENT, "split is 0 (sign) or 1 (median)");
      h->params.split = static_cast<int>(value);
    } else if (n == "taxicab_l1") {
      h->params.taxicab_l1 = value != 0;
    } else if (n == "scale_all") {
      h->params.scale_all = value != 0;
    } else if (n == "fixed") {
      // every fixed-mode semantic at once
      h->params.cosine_fixed = value != 0;
      h->params.taxicab_l1 = value != 0;
      h->params.scale_all = value != 0;
    } else if (n == "io_threads") {
      if (value < 0 || value > 256) return fail(h, AKNN_ERR_ARGUMENT, "io_threads must be in [0, 256]");
      h->io_threads = static_cast<int>(value);
    } else if (n == "pca_screen") {
      h->params.pca_screen = value != 0;
    } else if (n == "pca_fp16") {
      h->params.pca_fp16 = value != 0;
    } else if (n == "mean_fp64") {
      h->params.mean_fp64 = value != 0;
    } else if (n == "cosine_table") {
      h->params.cosine_table = value != 0;
    } else if (n == "graph_rounds") {
      h->params.graph_rounds = value != 0;
    } else if (n == "leaf_batch") {
      if (value < 0 || value > 2) return fail(h, AKNN_ERR_ARGUMENT, "leaf_batch is 0, 1 or 2");
      h->params.leaf_batch = static_cast<int>(value);
    } else if (n == "router_trees") {
      if (value < 0 || value > 1 << 20) return fail(h, AKNN_ERR_ARGUMENT, "router_trees must be >= 0");
      h->params.router_trees = static_cast<int>(value);
    } else {
      return fail(h, AKNN_ERR_ARGUMENT, "unknown option " + n);
    }
    // options apply to the next calculate(); the data stays uploaded
    if (h->engine) h->engine->set_options(h->params);
    return AKNN_OK;
  });
}

int aknn_get_graph(aknn_handle* h, int32_t* ids, float* vals) {
  return guarded(h, [&] {
    if (!h->engine || !h->calculated) return fail(h, AKNN_ERR_STATE, "nothing calculated");
    if (!ids || !vals) return fail(h, AKNN_ERR_ARGUMENT, "null output");
    h->engine->get_graph(ids, vals);
    return AKNN_OK;
  });
}

int aknn_get_normalized(aknn_handle* h, float* x) {
  return guarded(h, [&] {
    if (!h->engine || !h->calculated) return fail(h, AKNN_ERR_STATE, "nothing calculated");
    if (!x) return fail(h, AKNN_ERR_ARGUMENT, "null output");
    h->engine->get_normalized(x);
    return AKNN_OK;
  });
}

int aknn_get_forest(aknn_handle* h, int tree, int32_t* perm, int32_t* range_start, int32_t* range_count,
                    int32_t* iterations) {
  return guarded(h, [&] {
    if (!h->engine || !h->calculated) return fail(h, AKNN_ERR_STATE, "nothing calculated");
    if (!perm || !range_start || !range_count) return fail(h, AKNN_ERR_ARGUMENT, "null output");
    h->engine->get_forest(tree, perm, range_start, range_count);
    if (iterations) *iterations = h->engine->tree_iterations(tree);
    return AKNN_OK;
  });
}

int aknn_get_stage_times(aknn_handle* h, double* ms, int n) {
  return guarded(h, [&] {
    if (!h->engine || !h->calculated) return fail(h, AKNN_ERR_STATE, "nothing calculated");
    const aknn::StageTimes& t = h->engine->times();
    const double v[5] = {t.normalize, t.forest, t.initial, t.propagation, t.total};
    for (int i = 0; i < n && i < 5; ++i) ms[i] = v[i];
    return AKNN_OK;
  });
}

int aknn_get_kernel_stats(aknn_handle* h, const char* name, int64_t* launches, double* total_ms, double* units) {
  return guarded(h, [&] {
    if (!h->engine || !h->calculated) return fail(h, AKNN_ERR_STATE, "nothing calculated");
    if (!name) return fail(h, AKNN_ERR_ARGUMENT, "null name");
    const std::string n(name);
    const aknn::KernelStats* s = nullptr;
    if (n == "leaf_knn")
      s = &h->engine->stats_leaf_knn;
    else if (n == "propagate")
      s = &h->engine->stats_propagate;
    else if (n == "forest")
      s = &h->engine->stats_forest;
    else if (n == "propagate_graph") {
      // the rounds of the last calculate() replayed from the captured graph
      const bool gr = h->engine->rounds_from_graph();
      if (launches) *launches = gr ? h->engine->stats_propagate.launches : 0;
      if (total_ms) *total_ms = gr ? h->engine->stats_propagate.total_ms : 0.0;
      if (units) *units = gr ? h->engine->stats_propagate.units : 0.0;
      return AKNN_OK;
    } else if (n == "leaf_mfma") {
      // the screened (matrix-core) trees of the leaf k-NN: GEMM-form flops
      const aknn::KernelStats& l = h->engine->stats_leaf_knn;
      const int ldh = (h->engine->d() + 7) / 8 * 8;
      if (launches) *launches = l.screened_launches;
      if (total_ms) *total_ms = l.screened_ms;
      if (units) *units = 2.0 * ldh * l.screened;
      return AKNN_OK;
    } else if (n == "leaf_exact") {
      // the leaf k-NN trees that ran the exact FP32 chain on every pair
      // (all trees without the screen): 3D flops per pair
      const aknn::KernelStats& l = h->engine->stats_leaf_knn;
      const double per_pair = 3.0 * h->engine->d();
      if (launches) *launches = l.segments - l.screened_launches;
      if (total_ms) *total_ms = l.total_ms - l.screened_ms;
      if (units) *units = l.units - per_pair * l.screened;
      return AKNN_OK;
    } else
      return fail(h, AKNN_ERR_ARGUMENT, "unknown kernel " + n);
    if (launches) *launches = s->launches;
    if (total_ms) *total_ms = s->total_ms;
    if (units) *units = s->units;
    return AKNN_OK;
  });
}

int aknn_get_leaf_form(aknn_handle* h, int32_t* form, double* per_tree_ms, double* all_trees_ms) {
  return guarded(h, [&] {
    if (!h->engine || !h->calculated) return fail(h, AKNN_ERR_STATE, "nothing calculated");
    const int f = h->engine->leaf_form(per_tree_ms, all_trees_ms);
    if (form) *form = f;
    return AKNN_OK;
  });
}

int aknn_get_candidate_counts(aknn_handle* h, const char* stage, double* examined, double* rescored) {
  return guarded(h, [&] {
    if (!h->engine || !h->calculated) return fail(h, AKNN_ERR_STATE, "nothing calculated");
    if (!stage) return fail(h, AKNN_ERR_ARGUMENT, "null stage");
    const std::string n(stage);
    const aknn::KernelStats* s = nullptr;
    if (n == "cosine_table") {
      // the COSINE parity initial pass of the last build: 1 when its keys came
      // from the per-tree key table
      if (examined) *examined = h->engine->cosine_table_used() ? 1.0 : 0.0;
      if (rescored) *rescored = 0.0;
      return AKNN_OK;
    }
    if (n == "pca") {
      // the principal-subspace screen of the last build: its width (0: unused)
      // and the share of the sample's second moment in the top 16 directions
      if (examined) *examined = h->engine->pca_rank();
      if (rescored) *rescored = h->engine->pca_retained16();
      return AKNN_OK;
    }
    if (n == "propagate")
      s = &h->engine->stats_propagate;
    else if (n == "leaf_knn")
      s = &h->engine->stats_leaf_knn;
    else if (n == "forest")
      s = &h->engine->stats_forest;
    else
      return fail(h, AKNN_ERR_ARGUMENT, "unknown stage " + n);
    if (examined) *examined = n == "leaf_knn" ? s->screened : s->examined;
    if (rescored) *rescored = s->rescored;
    return AKNN_OK;
  });
}

int aknn_exact_knn(aknn_handle* h, const int64_t* queries, int64_t nq, int kq, int32_t* ids, double* dist) {
  return guarded(h, [&] {
    if (!h->engine || !h->calculated) return fail(h, AKNN_ERR_STATE, "nothing calculated");
    if (nq < 0 || (nq > 0 && (!queries || !ids || !dist))) return fail(h, AKNN_ERR_ARGUMENT, "null array");
    h->engine->exact_knn(queries, nq, kq, ids, dist);
    return AKNN_OK;
  });
}

int aknn_recall(aknn_handle* h, const int64_t* queries, int64_t nq, double* recall) {
  return guarded(h, [&] {
    if (!h->engine || !h->calculated) return fail(h, AKNN_ERR_STATE, "nothing calculated");
    if (nq < 0 || !recall || (nq > 0 && !queries)) return fail(h, AKNN_ERR_ARGUMENT, "null array");
    *recall = h->engine->recall(queries, nq);
    return AKNN_OK;
  });
}

int aknn_recall_noalloc(aknn_handle* h, const int64_t* queries, int64_t nq, double* recall) {
  return guarded(h, [&] {
    if (!h->engine || !h->calculated) return fail(h, AKNN_ERR_STATE, "nothing calculated");
    if (nq < 0 || !recall || (nq > 0 && !queries)) return fail(h, AKNN_ERR_ARGUMENT, "null array");
    *recall = aknn::recall_no_alloc(*h->engine, queries, nq);
    return AKNN_OK;
  });
}

int aknn_set_bins(aknn_handle* h, const int32_t* bins, int64_t n, int bin_bits) {
  return guarded(h, [&] {
    if (!h->engine) return fail(h, AKNN_ERR_STATE, "not initialized");
    if (!bins) return fail(h, AKNN_ERR_ARGUMENT, "null bins");
    if (n != h->engine->n()) return fail(h, AKNN_ERR_ARGUMENT, "bins must hold one entry per point");
    h->engine->set_bins(bins, bin_bits);
    return AKNN_OK;
  });
}

int aknn_read_points(const char* path, int64_t n, int d, int klaster, int threads, float* x, int32_t* labels,
                     char* error, int error_len) {
  std::string err;
  try {
    if (!path || !x || n < 1 || d < 1 || (klaster && !labels) || threads < 0) {
      err = "bad arguments";
    } else {
      std::vector<float> xs;
      std::vector<int32_t> ls;
      const std::string p(path);
      err = aknn::is_fbin(p) ? aknn::load_fbin(p, n, d, xs)
                             : (threads == 1 ? aknn::load_csv(p, n, d, klaster != 0, xs, ls)
                                             : aknn::load_csv_parallel(p, n, d, klaster != 0, threads, xs, ls));
      if (err.empty()) {
        std::memcpy(x, xs.data(), sizeof(float) * xs.size());
        if (klaster && !aknn::is_fbin(p)) std::memcpy(labels, ls.data(), sizeof(int32_t) * ls.size());
      }
    }
  } catch (const std::exception& e) {
    err = e.what();
  } catch (...) {
    err = "unknown error";
  }
  if (error && error_len > 0) {
    std::strncpy(error, err.c_str(), static_cast<size_t>(error_len) - 1);
    error[error_len - 1] = '\0';
  }
  return err.empty() ? AKNN_OK : AKNN_ERR_IO;
}

aknn_router* aknn_router_create(int device, int d, int hidden, int bin_bits, uint32_t seed) {
  try {
    auto* r = new aknn_router();
    r->router = std::make_unique<aknn::Router>(device, d, hidden, bin_bits, seed);
    return r;
  } catch (...) {
    return nullptr;
  }
}

void aknn_router_destroy(aknn_router* r) {
  try {
    delete r;
  } catch (...) {
  }
}

int aknn_router_labels(aknn_router* r, aknn_handle* h, int sweeps, int32_t* labels, double* cut) {
  return rguarded(r, [&] {
    if (!h || !h->engine || !h->calculated) return rfail(r, AKNN_ERR_STATE, "the graph handle has no graph");
    if (!labels || sweeps < 0) return rfail(r, AKNN_ERR_ARGUMENT, "bad arguments");
    std::vector<int32_t> lab;
    r->router->graph_labels(*h->engine, sweeps, lab);
    std::memcpy(labels, lab.data(), sizeof(int32_t) * lab.size());
    if (cut) {
      cut[0] = r->router->train_stats().cut_before;
      cut[1] = r->router->train_stats().cut_after;
    }
    return AKNN_OK;
  });
}

int aknn_router_fit(aknn_router* r, aknn_handle* h, const int32_t* labels, int sweeps, int epochs, int batch,
                    float lr) {
  return rguarded(r, [&] {
    if (!h || !h->engine || !h->calculated) return rfail(r, AKNN_ERR_STATE, "the graph handle has no graph");
    std::vector<int32_t> lab;
    if (labels)
      lab.assign(labels, labels + h->engine->n());
    else
      r->router->graph_labels(*h->engine, sweeps, lab);
    r->router->fit(*h->engine, lab, epochs, batch, lr);
    return AKNN_OK;
  });
}

int aknn_router_gradients(aknn_router* r, const float* x, int64_t m, const int32_t* labels, float* grads,
                          double* loss) {
  return rguarded(r, [&] {
    if (!x || !labels || !grads) return rfail(r, AKNN_ERR_ARGUMENT, "null array");
    const double l = r->router->gradients(x, m, labels, grads);
    if (loss) *loss = l;
    return AKNN_OK;
  });
}

int aknn_router_logits(aknn_router* r, const float* x, int64_t m, float* z) {
  return rguarded(r, [&] {
    if (!x || !z) return rfail(r, AKNN_ERR_ARGUMENT, "null array");
    r->router->logits(x, m, z);
    return AKNN_OK;
  });
}

int aknn_router_bins(aknn_router* r, const float* x, int64_t m, int fused, int32_t* bins) {
  return rguarded(r, [&] {
    if (!x || !bins) return rfail(r, AKNN_ERR_ARGUMENT, "null array");
    r->router->set_fused(fused != 0);
    try {
      r->router->bins_of(x, m, bins);
    } catch (...) {
      r->router->set_fused(true);
      throw;
    }
    r->router->set_fused(true);
    return AKNN_OK;
  });
}

int aknn_router_param_count(aknn_router* r, int64_t* count) {
  return rguarded(r, [&] {
    if (!count) return rfail(r, AKNN_ERR_ARGUMENT, "null output");
    *count = static_cast<int64_t>(r->router->param_count());
    return AKNN_OK;
  });
}

int aknn_router_get_params(aknn_router* r, float* params, float* mu, float* inv_sigma) {
  return rguarded(r, [&] {
    r->router->get_params(params, mu, inv_sigma);
    return AKNN_OK;
  });
}

int aknn_router_set_params(aknn_router* r, const float* params, const float* mu, float inv_sigma) {
  return rguarded(r, [&] {
    r->router->set_params(params, mu, inv_sigma);
    return AKNN_OK;
  });
}

int aknn_router_save(aknn_router* r, const char* path) {
  return rguarded(r, [&] {
    if (!path) return rfail(r, AKNN_ERR_ARGUMENT, "null path");
    try {
      r->router->save(path);
    } catch (const std::runtime_error& e) {
      return rfail(r, AKNN_ERR_IO, e.what());
    }
    return AKNN_OK;
  });
}

int aknn_router_load(aknn_router* r, const char* path) {
  return rguarded(r, [&] {
    if (!path) return rfail(r, AKNN_ERR_ARGUMENT, "null path");
    try {
      r->router->load(path);
    } catch (const std::runtime_error& e) {
      return rfail(r, AKNN_ERR_IO, e.what());
    }
    return AKNN_OK;
  });
}

int aknn_router_stats(aknn_router* r, double* out, int n) {
  return rguarded(r, [&] {
    if (!out) return rfail(r, AKNN_ERR_ARGUMENT, "null output");
    const aknn::RouterTrainStats& t = r->router->train_stats();
    const double v[7] = {static_cast<double>(t.epochs), t.loss, t.accuracy, t.train_ms, t.cut_before, t.cut_after,
                         r->router->last_assign_ms()};
    for (int i = 0; i < n && i < 7; ++i) out[i] = v[i];
    return AKNN_OK;
  });
}

const char* aknn_router_last_error(aknn_router* r) { return r ? r->error.c_str() : "null router"; }

int aknn_set_router(aknn_handle* h, aknn_router* r) {
  return guarded(h, [&] {
    if (!h->engine) return fail(h, AKNN_ERR_STATE, "not initialized");
    if (r && !r->router) return fail(h, AKNN_ERR_ARGUMENT, "empty router");
    if (r && (r->router->device() != h->device || r->router->d() != h->engine->d()))
      return fail(h, AKNN_ERR_ARGUMENT, "the router must have the data's D and live on the handle's device");
    h->engine->set_router(r ? r->router.get() : nullptr);
    return AKNN_OK;
  });
}

int aknn_add_router(aknn_handle* h, aknn_router* r) {
  return guarded(h, [&] {
    if (!h->engine) return fail(h, AKNN_ERR_STATE, "not initialized");
    if (!r || !r->router) return fail(h, AKNN_ERR_ARGUMENT, "empty router");
    if (r->router->device() != h->device || r->router->d() != h->engine->d())
      return fail(h, AKNN_ERR_ARGUMENT, "the router must have the data's D and live on the handle's device");
    h->engine->add_router(r->router.get());
    return AKNN_OK;
  });
}

const char* aknn_last_error(aknn_handle* h) { return h ? h->error.c_str() : "null handle"; }

}  // extern "C"
