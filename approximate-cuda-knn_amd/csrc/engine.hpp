// The k-NN graph engine: device buffers and the calculate() bracket of
// EuclideanDistanceMatrixGPU (euclideandistancematrixgpu.cu:2861-2896).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "aknn_common.hpp"
#include "glibc_rand.hpp"
#include "kernels.hpp"

namespace aknn {

struct Params {
  int k = 0;
  int trees = 0;
  int leaf = 0;   // minSize: a bucket with more points than this is split
  int props = 0;
  int metric = kEuclidean;
  uint32_t seed = 1;      // rand() stream seed (the reference never seeds: 1)
  int rng_order = 0;      // 0: Leaf7 gets (second, first) draw of a level; 1: (first, second)
  int cosine_fixed = 0;   // 1: carry COSINE distances instead of the 1 - key flip
  int level_batch = 16;   // forest levels enqueued per host check
  int skip_settled = 1;   // propagation skips neighbours whose rows were examined last round (same results)
  int sketch = 0;         // propagation screens candidates with the FP16 sketch bound (same results)
  int leaf_mfma = 0;      // leaf k-NN of trees after the first screens pairs on the matrix cores (same results)
  int sketch_int8 = 0;    // with sketch, EUCLIDEAN propagation screens with the int8 sketch (same results)
  int forest_int8 = 1;    // with sketch, the forest side test of staged buckets reads the int8 sketch (same results)
  int flip_int8 = 1;      // COSINE parity rounds screen with the int8 sketch where it is made (same results)
  int split = 0;          // forest rule: 0 the reference's sign split, 1 median split on seeded directions
  int router_trees = 0;   // median split: trees 0..router_trees-1 take the router bins as their first levels
  int taxicab_l1 = 0;     // TAXICAB keys and stores the L1 total instead of its square root
  int scale_all = 0;      // normalization scales every dimension, not only the first 256
  int graph_rounds = 1;   // once the round mode is fixed, replay the propagation rounds as one captured hipGraph
  int pca_screen = 0;     // EUCLIDEAN propagation screens with a principal-subspace bound when the data allow (same results)
  int pca_fp16 = 1;       // the screen's projections are kept in FP16, their rounding inside the radii (same results)
  int cosine_table = 1;   // COSINE parity initial pass reads its keys from a per-tree key table when it fits (same results)
  int mean_fp64 = 0;      // normalization's mean as blocked FP64 sums (the reference: a sequential FP32 sum)
  int leaf_batch = 2;     // initial leaf k-NN launches: 0 one per tree, 1 all trees in one persistent launch,
                          // 2 the faster of the two, timed on the first calculate() of a data set (same results)
};

// Per-stage times of the last calculate(), milliseconds.
struct StageTimes {
  double normalize = 0, forest = 0, initial = 0, propagation = 0, total = 0;
};

// Launch-level statistics of one kernel family in the last calculate().
struct KernelStats {
  int64_t launches = 0;
  double total_ms = 0;
  double units = 0;  // algorithmic bytes or flops summed over launches
  double examined = 0;  // propagation: new candidates examined
  double rescored = 0;  // propagation: candidates whose exact distance was computed;
                        // leaf k-NN: pairs of screened trees whose exact distance was computed
  double screened = 0;  // leaf k-NN: pairs screened on the matrix cores
  int64_t screened_launches = 0;  // leaf k-NN: trees that ran screened
  double screened_ms = 0;
  int64_t segments = 0;  // leaf k-NN: trees (one launch may run all of them)
};

class Engine {
 public:
  explicit Engine(int device);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  void set_data(const float* x_rowmajor, int64_t n, int d, const Params& p);
  void calculate();
  // seed, rng_order, cosine_fixed, level_batch of p (shape parameters are fixed by set_data)
  void set_options(const Params& p) {
    p_.seed = p.seed;
    p_.rng_order = p.rng_order;
    p_.cosine_fixed = p.cosine_fixed;
    p_.level_batch = p.level_batch;
    p_.skip_settled = p.skip_settled;
    if (p_.sketch != p.sketch || p_.sketch_int8 != p.sketch_int8) prop_mode_ = -1;
    p_.sketch = p.sketch;
    p_.leaf_mfma = p.leaf_mfma;
    p_.sketch_int8 = p.sketch_int8;
    p_.forest_int8 = p.forest_int8;
    if (p_.flip_int8 != p.flip_int8) prop_mode_ = -1;
    p_.flip_int8 = p.flip_int8;
    if (p_.split != p.split) prop_mode_ = -1;
    p_.split = p.split;
    p_.router_trees = p.router_trees;
    if (p_.taxicab_l1 != p.taxicab_l1 || p_.scale_all != p.scale_all) prop_mode_ = -1;
    p_.taxicab_l1 = p.taxicab_l1;
    p_.scale_all = p.scale_all;
    p_.graph_rounds = p.graph_rounds;
    p_.pca_screen = p.pca_screen;
    if (p_.pca_fp16 != p.pca_fp16) prop_mode_ = -1;
    p_.pca_fp16 = p.pca_fp16;
    p_.cosine_table = p.cosine_table;
    if (p_.mean_fp64 != p.mean_fp64) prop_mode_ = -1;
    p_.mean_fp64 = p.mean_fp64;
    p_.leaf_batch = p.leaf_batch;
    leaf_form_ = -1;  // the leaf launch form is learned again under the new options
    drop_round_graph();
  }
  // Router bins of the points (host array [n], values < 2^bin_bits): the
  // first bin_bits levels of routed median-split trees split on their bits.
  void set_bins(const int32_t* bins, int bin_bits);
  // Bins already on the device (router inference), same contract.
  void set_bins_device(const int32_t* bins, int bin_bits);
  int32_t* bins_device() { return bins_; }
  // Trained routers that assign the bins of routed trees inside calculate()
  // (routed tree t takes router t mod R); they must live on this engine's
  // device.  set_router replaces the list (nullptr: empty), add_router appends.
  void set_router(class Router* r) {
    routers_.clear();
    if (r) routers_.push_back(r);
  }
  void add_router(class Router* r) {
    if (r) routers_.push_back(r);
  }
  const float* raw_points() const { return x_raw_; }
  int ld() const { return ld_; }
  // normalized points of the last calculate(), the current graph's ids, and
  // the stream every launch of this engine goes to
  const float* points() const { return x_; }
  const int32_t* graph_ids() const { return gids_[cur_]; }
  hipStream_t stream() const { return stream_; }

  void get_graph(int32_t* ids, float* vals) const;
  void get_normalized(float* x) const;
  void get_forest(int tree, int32_t* perm, int32_t* rs, int32_t* rc);
  // FP64 exact k-NN of query points over the normalized points (host arrays)
  void exact_knn(const int64_t* queries, int64_t nq, int kq, int32_t* ids, double* dist);
  // recall@k of the current graph against the FP64 exact k-NN of the queries
  double recall(const int64_t* queries, int64_t nq);
  bool rounds_from_graph() const { return rounds_from_graph_; }
  // basis width of the last build's principal-subspace screen (0: not used)
  int pca_rank() const { return pca_.r; }
  bool cosine_table_used() const { return cosine_table_used_; }
  double pca_retained16() const { return pca_.retained16; }
  // leaf launch form of the last calculate() (0 per tree, 1 all trees) and
  // the times the learning call measured for each (0 until it ran)
  int leaf_form(double* per_tree_ms, double* all_trees_ms) const {
    if (per_tree_ms) *per_tree_ms = leaf_tune_ms_[0];
    if (all_trees_ms) *all_trees_ms = leaf_tune_ms_[1];
    return leaf_all_used_ ? 1 : 0;
  }
  int tree_iterations(int tree) const { return tree < static_cast<int>(iters_.size()) ? iters_[tree] : -1; }

  const StageTimes& times() const { return times_; }
  KernelStats stats_leaf_knn, stats_propagate, stats_forest;

  int64_t n() const { return n_; }
  int d() const { return d_; }
  const Params& params() const { return p_; }
  int device() const { return device_; }

 private:
  void release();
  template <typename T>
  T* alloc(size_t count);
  // propagation screens with the sketch (not in COSINE parity mode, whose
  // flipped stored values give no distance threshold)
  // metric code of the distance kernels
  int kmetric() const { return p_.metric == kTaxicab && p_.taxicab_l1 ? kTaxicabL1 : p_.metric; }
  bool propagation_screened() const { return p_.sketch != 0; }
  bool leaf_mfma_active() const {
    return p_.leaf_mfma && p_.trees > 1 && (p_.metric == kEuclidean || (p_.metric == kCosine && p_.cosine_fixed));
  }
  void build_forest();
  void build_forest_median();
  // the propagation rounds as one hipGraph (captured once the round mode is
  // fixed; every option change drops it)
  void drop_round_graph();
  bool propagate_graph(int base);
  void initial_knn();
  void propagate();
  void round_counts(int r, unsigned long long out[2]);
  void slot_sums(const unsigned long long* dev, int kinds, unsigned long long* out);

  int device_;
  hipStream_t stream_ = nullptr;
  std::vector<void*> allocs_;
  std::vector<hipEvent_t> events_;
  hipEvent_t ev(int i);
  hipEvent_t launch_ev(int64_t i);
  std::vector<hipEvent_t> launch_events_;

  Params p_;
  int64_t n_ = 0;
  int d_ = 0, ld_ = 0;
  int leaf_eff_ = 1;
  float* x_ = nullptr;
  float* x_raw_ = nullptr;  // pristine copy: calculate() normalizes x_ in place, re-runs start from here
  float* norms_ = nullptr;
  _Float16* xh_ = nullptr;  // FP16 sketch [n][ldh_], allocated on first use
  float* eps_ = nullptr;    // sketch radii [n]
  float* nh_ = nullptr;     // sketch squared norms [n]
  int8_t* q8_ = nullptr;    // int8 sketch [n][ld8_], scales, sums of squares, radii
  float* s8_ = nullptr;
  int32_t* a8_ = nullptr;
  float* e8_ = nullptr;
  float2* se8_ = nullptr;  // (scale, radius) pairs of the int8 sketch for the forest side test [n]
  bool q8_ready_ = false;   // the int8 sketch was made by the current build
  int ld8_ = 0;
  unsigned long long* leaf_rescored_ = nullptr;
  int ldh_ = 0;
  float* mean_ = nullptr;
  double* mean64_ = nullptr;  // blocked FP64 partial sums of the mean (option mean_fp64), allocated on first use
  uint32_t* gmax_ = nullptr;
  int32_t* gids_[2] = {nullptr, nullptr};
  float* gvals_[2] = {nullptr, nullptr};
  int cur_ = 0;
  int32_t* perm_ = nullptr;  // [trees][n]
  Leaf* leaves_ = nullptr;   // [trees][n]
  std::vector<int> nleaves_, iters_;
  ForestBuffers fb_{};
  ForestCtl* ctl_host_ = nullptr;
  unsigned long long* ncand_ = nullptr;
  uint8_t* changed_[2] = {nullptr, nullptr};
  int32_t* rs_ = nullptr;  // [trees][n] per-point start ranges, allocated on demand
  Leaf* items_ = nullptr;  // leaf k-NN work items of all trees, tree after tree
  size_t items_cap_ = 0;   // capacity of items_ (entries)
  std::vector<std::vector<Leaf>> items_host_;
  // pinned host staging, grown on demand: 0 leaf lists, 1 work items,
  // 2 leaf-screen counters after each tree, 3 the all-trees launch's tree
  // modes, clock stamps and error flag
  struct Pinned {
    void* p = nullptr;
    size_t cap = 0;
  };
  Pinned pinned_[5];
  void* pinned(int slot, size_t bytes);
  int32_t* rc_ = nullptr;
  // the all-trees leaf launch's state (per-tree arguments, queues, barrier,
  // modes, stamps, counters, error), allocated on first use
  LeafAllState leaf_st_{};
  bool leaf_all_used_ = false;
  int leaf_form_ = -1;  // leaf launch form learned on this data set: -1 unknown, 0 per tree, 1 all trees
  double leaf_tune_ms_[2] = {0, 0};  // the learning call's times of the two forms
  int wall_khz_ = 0;
  bool initial_knn_all(const Leaf* lh, const std::vector<int64_t>& leaf_off, bool mfma);
  void initial_knn_big();
  void propagate_big();
  void leaf_all_stats(bool mfma);
  CosineParityScratch cps_{};
  // COSINE parity key table: offsets [trees][n] (+ total), block sums of
  // their scan, the keys, and whether the last build used it
  int64_t* koff_ = nullptr;
  int64_t* kbsum_ = nullptr;
  float* ktab_ = nullptr;
  int64_t ktab_cap_ = 0;
  bool cosine_table_used_ = false;
  int64_t active_sum_ = 0;
  std::vector<uint8_t> round_screened_, tree_screened_;  // which rounds / trees used the screens
  std::vector<uint8_t> round_form_;  // per round the screen form launched (launch_propagate's return)
  std::vector<uint8_t> graph_form_;  // the same for the rounds captured in round_exec_
  int prop_mode_ = -1;  // propagation rounds learned on this data: -1 unknown, 0 screened, 1 exact
  DrawStream draws_;
  // median-split forest: directions [trees][nd][ld] (cached per seed),
  // projections of a batch of trees [batch][nd][n], router bins [n]; the
  // batch's level-loop buffers (fbm_, for fbm_trees_ trees at once) and
  // per-tree leaf counts and last split levels (tree_ctl_ [2][fbm_trees_])
  float* dirs_ = nullptr;
  float* proj_ = nullptr;
  ForestBuffers fbm_{};
  int fbm_trees_ = 0;
  int proj_nd_ = 0;
  int32_t* tree_ctl_ = nullptr;
  int median_batch_trees(int nd);
  int nd_ = 0;
  int64_t dirs_key_ = -1;
  int64_t proj_dirs_ = 0;      // directions projected in the last median forest
  int64_t proj_launches_ = 0;  // project_kernel launches (passes over the rows)
  int32_t* bins_ = nullptr;
  int bin_bits_ = 0;
  int bins_cap_ = 0;  // bin arrays bins_ holds
  bool bins_set_ = false;
  std::vector<class Router*> routers_;
  // principal-subspace screen: basis [32][d] and moments scratch [d][d],
  // projections [n][32] and radii [n]; pca_ is this build's basis
  float* pca_p_ = nullptr;
  float* pca_mom_ = nullptr;
  float* pca_proj_ = nullptr;
  float* pca_e_ = nullptr;
  PcaBasis pca_;
  hipEvent_t pca_ev_ = nullptr;  // the sample moments have reached the host
  hipStream_t aux_ = nullptr;     // side stream (default priority): router inference beside normalization
  hipEvent_t router_ev_ = nullptr;
  bool routers_launched_ = false;  // this build's router bins are being made on aux_
  hipGraphExec_t round_exec_ = nullptr;
  int round_graph_mode_ = -1;
  double round_graph_sigma_ = 0.0;
  bool rounds_from_graph_ = false;
  StageTimes times_;
};

// recall@k of e's current graph over host query indices, computed with the
// kernels of Engine::recall on buffers carved from a static device array
// (no device allocation); rejects queries whose buffers exceed that array.
// Defined in noalloc.hip.
double recall_no_alloc(Engine& e, const int64_t* queries, int64_t nq);
// bytes of that array a call with these shapes uses
size_t recall_scratch_bytes(int64_t n, int k, int64_t nq, bool cosine);

}  // namespace aknn
