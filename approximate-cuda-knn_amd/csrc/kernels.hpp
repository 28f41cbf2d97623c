// Host-side launchers of the HIP kernels (one translation unit per stage).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "aknn_common.hpp"

namespace aknn {

// ---- normalization (euclideandistancematrixgpu.cu:19-85) -------------------
// x: [n][ld] row-major, d <= ld.  mean: [d] scratch, gmax: [1] scratch.
// scale_all: every dimension enters the max-abs and is scaled (the reference
// does so for the first 256 only).  mean64_part (nullable, ceil(n / 2048) * d
// doubles): the mean as blocked FP64 sums instead of the reference's
// sequential FP32 sum.
//
// The FP16 sketch outputs of launch_normalize's row pass (device_util.hpp,
// sketch_element): xh [n][ldh] halves (ldh % 8 == 0, ldh >= d); eps[n]: upper
// bounds of the rounding error's norm (l1: L1 for TAXICAB, else L2); nh[n]:
// ||xh||^2.
struct SketchOut {
  bool l1 = false;
  _Float16* xh = nullptr;
  int ldh = 0;
  float* eps = nullptr;
  float* nh = nullptr;
};
// src: the uploaded points (read only); x: the normalized points (every
// element of every row written, the padding copied).  sk (nullable): the
// FP16 sketch of the normalized rows, formed in the same pass.
void launch_normalize(const float* src, float* x, int64_t n, int d, int ld, float* mean, uint32_t* gmax,
                      bool scale_all, double* mean64_part, const SketchOut* sk, hipStream_t s);
// Per-point COSINE squared norms (the chain of .cu:849-865 for one operand).
void launch_norms(const float* x, int64_t n, int d, int ld, float* norms, hipStream_t s);

// ---- forest (euclideandistancematrixgpu.cu:2130-2686) -----------------------
struct ForestBuffers {
  int32_t* act_ids[2];
  int32_t* act_bkt[2];
  Bucket* bkt[2];
  uint64_t* rmask;
  int32_t* wpre;
  int32_t* blk_r;
  int32_t* blk_roff;
  // per bucket of a level: children, then (after the scan) the index of the
  // first child; active elements of the children, then their first offset
  int32_t* bchild;
  int32_t* bact;
  // per bucket of a level (even/odd): the ids of its two pivots, written by
  // the elements that land on them (nullable: the side test looks them up)
  int32_t* piv[2] = {nullptr, nullptr};
  ForestCtl* ctl;
  int64_t n;
  int64_t max_buckets;
  // FP16 sketch for the side test (nullable: exact chains for every point;
  // the partition is the same), and a counter of points it left undecided
  // (kSlotWords u64, spread over kCountSlots slots)
  const _Float16* xh = nullptr;
  int ldh = 0;
  const float* eps = nullptr;
  const float* nh = nullptr;
  unsigned long long* amb_count = nullptr;
  // int8 sketch for the side test of buckets staged in LDS (nullable: the
  // FP16 sketch decides): rows [n][ld8], (scale, radius) pairs
  const int8_t* q8 = nullptr;
  int ld8 = 0;
  const float2* se8 = nullptr;
};
void launch_forest_init(const ForestBuffers& fb, int32_t* perm, Leaf* leaves, int64_t n, int32_t q1, int32_t q2,
                        int leaf_eff, int kp1, hipStream_t s);
void launch_forest_level(const ForestBuffers& fb, const float* x, int d, int ld, int32_t* perm, Leaf* leaves,
                         int level, int32_t v1, int32_t v2, int leaf_eff, int kp1, hipStream_t s);
// Median-split forest.  launch_project: proj[j * n + i] = x_i . dirs[j] (FP32
// chain in dimension order), j < nd <= kProjectStep, dirs [nd][ld], one pass
// over all rows in order (fb null) or over the rows of the points active at
// `level` (fb's list for its parity).
constexpr int kProjectStep = 16;
void launch_project(const float* x, int64_t n, int d, int ld, const float* dirs, int nd, float* proj,
                    const ForestBuffers* fb, int level, hipStream_t s);
// Up to kProjSetMax groups of directions (several trees' first directions)
// projected in one pass over all rows in order: group q writes
// proj[q][j * n + i] = x_i . dirs[q][j] for j < nd[q] <= kProjectStep.
constexpr int kProjSetMax = 6;
struct ProjSet {
  const float* dirs[kProjSetMax];
  float* proj[kProjSetMax];
  int nd[kProjSetMax];
};
void launch_project_set(const float* x, int64_t n, int d, int ld, const ProjSet& ps, int groups, hipStream_t s);
// A batch of median-split trees is built level by level together: one
// active list, one bucket list (each bucket records its batch tree) and the
// same three passes per level for every tree.  Batch tree tl (global index
// tree0 + tl) splits on bit bb - 1 - level of its router bins for its first
// bb levels (bb = bin_bits when tree0 + tl < routed_trees, else 0), then on
// the sampled median of its projections on direction level - bb (proj [batch
// trees][nd][n]), then by position.
struct MedianRules {
  int tree0 = 0;
  int routed_trees = 0;
  int bin_bits = 0;
  int nsets = 1;                  // bin arrays: tree t reads bins + (t % nsets) * n
  const int32_t* bins = nullptr;  // [nsets][n]
  const float* proj = nullptr;
  int nd = 0;
  int64_t n = 0;
};
// fb sized for ntrees * n active points; perm, leaves: [ntrees][n];
// tree_nleaves, tree_last: [ntrees] (leaves written, last level that still
// split a bucket of the tree, -1 for none).
void launch_forest_init_batch(const ForestBuffers& fb, int ntrees, int32_t* perm, Leaf* leaves, int32_t* tree_nleaves,
                              int32_t* tree_last, int64_t n, int leaf_eff, hipStream_t s);
// any_median: some tree of the batch is at a projection level (its buckets
// need their sampled pivots).
void launch_forest_level_batch(const ForestBuffers& fb, const MedianRules& rules, bool any_median, int level,
                               int32_t* perm, Leaf* leaves, int32_t* tree_nleaves, int32_t* tree_last, int leaf_eff,
                               int kp1, hipStream_t s);
// Per-point start ranges from a leaf list (for export and the COSINE parity path).
void launch_leaf_ranges(const int32_t* perm, const Leaf* leaves, int nleaves, int32_t* rs, int32_t* rc,
                        hipStream_t s);

// ---- graph -------------------------------------------------------------------
void launch_stub_init(int32_t* ids, float* vals, int64_t count, hipStream_t s);

// Initial leaf k-NN for one tree, merged into the graph rows in place
// (EUCLIDEAN / TAXICAB / COSINE-fixed; exact top-k by (key, id)).  leaves:
// work items {first owner position, owners, candidate range start, count};
// with rsp/rcp (per-point start-node ranges, nullable) an item may group the
// owners of several leaves, each pair masked by its owner's own range.
void launch_leaf_knn(int metric, const float* x, int d, int ld, const float* norms, const int32_t* perm,
                     const Leaf* leaves, int nleaves, int max_leaf, int k, int32_t* ids, float* vals,
                     const int32_t* rsp, const int32_t* rcp, hipStream_t s);

// What the previous round left behind, so a round can skip neighbours whose
// rows were already fully examined (results are unchanged by the skip).
// ids_prev: the previous round's input graph (it is the buffer this round
// writes; nullable: no skipping); changed_prev[j]: row j changed in the
// previous round; changed_out[i]: written by this round (nullable).
struct PropagateHistory {
  const int32_t* ids_prev = nullptr;
  const uint8_t* changed_prev = nullptr;
  uint8_t* changed_out = nullptr;
};

// Candidate counters of one propagation round: two kinds (examined,
// re-scored), each spread over kCountSlots slots one 128-byte line apart so
// that millions of per-point additions do not queue on one address; the total
// is the sum of the slots.
constexpr int kCountSlots = 64;
constexpr int kCountStride = 16;                                  // u64 per slot (128 bytes)
constexpr int kSlotWords = kCountSlots * kCountStride;           // one spread counter
constexpr int kCountWords = 2 * kSlotWords;                       // per round

// One Jacobi propagation round, ids_in/vals_in -> ids_out/vals_out.
// flip: the COSINE parity store rule (stored = 1 - key; no skipping then,
// since any candidate triggers the flip).  order: processing order of the
// points (nullable: 0..n-1); results do not depend on it.  xh/eps: FP16
// sketch and radii (nullable: exact distances for every candidate; results do
// not depend on it).  ncand: optional device counters of the round
// (kCountWords u64: examined new candidates, exactly re-scored candidates;
// byte accounting).
struct PropagateArgs {
  int metric = kEuclidean;
  bool flip = false;
  const float* x = nullptr;
  int d = 0, ld = 0;
  const _Float16* xh = nullptr;
  int ldh = 0;
  const float* eps = nullptr;
  // int8 sketch (EUCLIDEAN; used instead of xh when set)
  const int8_t* q8 = nullptr;
  int ld8 = 0;
  const float* s8 = nullptr;
  const int32_t* a8 = nullptr;
  const float* e8 = nullptr;
  const float* norms = nullptr;
  // principal-subspace screen (EUCLIDEAN; used instead of the sketches when set)
  const float* pca = nullptr;
  const _Float16* pca16 = nullptr;  // FP16 projections (taken instead of pca when set)
  int pca_r = 0;
  const float* pca_e = nullptr;
  double pca_sigma = 1.0;
  const int32_t* order = nullptr;
  int64_t n = 0;
  int k = 0;
  const int32_t* ids_in = nullptr;
  const float* vals_in = nullptr;
  int32_t* ids_out = nullptr;
  float* vals_out = nullptr;
  unsigned long long* ncand = nullptr;
  PropagateHistory hist;
};
// Returns the screen form launched: 0 exact, 1 FP16 sketch, 2 int8 sketch,
// 3 FP32 and 4 FP16 principal-subspace projections.
int launch_propagate(const PropagateArgs& a, hipStream_t s);
// whether a propagation round with k neighbours over d dimensions fits the
// LDS of at least one wave (the exact pass; screens that do not fit give way
// to it)
bool propagation_fits(int k, int d);

// ---- principal-subspace screen (pca.hip) ------------------------------------
// r (16 or 32, 0: the data is not close enough to a 16- or 32-dimensional
// subspace, or D is outside [16, 256]) FP32 rows p [r][d] of top sample
// eigenvectors, sigma >= the largest singular value of p.
struct PcaBasis {
  int r = 0;
  std::vector<float> p;
  double sigma = 1.0;
  double retained16 = 0.0;  // share of the sample's second moment in the top 16 directions
};
// The sample's second moments into moments_dev[0, d*d) (device scratch of
// 33*d*d floats); false (nothing launched) when D is outside [16, 256] or
// n < 64.  pca_basis_from_moments: the basis from those moments on the host
// (subspace iteration and Rayleigh-Ritz in FP64; no device call).
bool launch_pca_moments(const float* x, int64_t n, int d, int ld, float* moments_dev, hipStream_t s);
PcaBasis pca_basis_from_moments(const float* moments, int d);
// proj [n][r] = fl(p x_i), rad[n] >= ||proj_i - p x_i||.  With proj16
// (proj then null): proj16 [n][r] = the FP16 roundings of fl(p x_i) and
// rad[n] >= ||proj16_i - p x_i|| (the rounding's exact norm added; infinite
// when a coordinate is beyond FP16's range).
void launch_pca_project(const float* x, int64_t n, int d, int ld, const float* p, int r, float* proj, float* rad,
                        hipStream_t s, _Float16* proj16 = nullptr);
// *out = max of n non-negative floats (+inf allowed).  The screens' radius
// arrays keep this at index n: one upper bound for every candidate, read
// once per wave instead of one scattered 4-byte load per candidate.
void launch_max_nonneg(const float* v, int64_t n, float* out, hipStream_t s);

// ---- sketches of the normalized points ---------------------------------------
// (the FP16 sketch: launch_normalize's SketchOut)
// Int8 sketch: q8 [n][ld8] (ld8 % 16 == 0, ld8 >= d), per-row scale s8,
// a8 = sum q^2, e8 = upper bound of ||x - s8 q8||_2; se (nullable): the
// pairs (s8, e8) side by side, one 8-byte load per point.
void launch_sketch8(const float* x, int64_t n, int d, int ld, int8_t* q8, int ld8, float* s8, int32_t* a8, float* e8,
                    float2* se, hipStream_t s);

// Initial leaf k-NN for one tree, EUCLIDEAN or COSINE with carried distances,
// with the candidate pairs screened by half-precision MFMA dot products of the
// sketch and proven lower bounds; the same graph as launch_leaf_knn.
// norms: the COSINE squared-norm chains; eps_max: device address of the
// largest sketch radius.  rescored, screened: optional device counters of
// pairs whose exact distance was computed and of pairs screened.
void launch_leaf_mfma(int metric, const float* x, int d, int ld, const _Float16* xh, int ldh, const float* eps,
                      const float* eps_max, const float* nh, const float* norms, const int32_t* perm, const Leaf* leaves, int nleaves,
                      int max_leaf, int k, int32_t* ids, float* vals, const int32_t* rsp, const int32_t* rcp,
                      unsigned long long* rescored, unsigned long long* screened, hipStream_t s);

// Every tree's initial leaf k-NN in one persistent cooperative launch (trees
// in order, a grid barrier between them; each tree's work items from eight
// per-XCD queues).  Tree 0 runs the exact body; with `screen` (EUCLIDEAN or
// COSINE-fixed and the sketch given) later trees run the MFMA-screened body
// until a screened tree re-scores more than 30% of its pairs.  Same graph as
// launch_leaf_knn / launch_leaf_mfma tree by tree.  perm, rs, rc: [trees][n];
// items: every tree's work items (device), tree after tree, at host offsets
// item_off [trees + 1].  Returns false, launching nothing, when the runtime
// refuses the cooperative grid.
struct LeafAllState {
  void* targs;                  // leaf_all_targs_bytes(trees): per-tree arguments
  unsigned* queue;              // [trees][8] dequeue counters
  unsigned* bar;                // [2] barrier arrivals and generation
  int32_t* mode;                // [trees] out: 1 where the tree ran screened
  unsigned long long* stamps;   // [trees + 1] out: wall clock at each tree's start and at the end
  unsigned long long* prev;     // [2] counter totals after the last screened tree
  int* error;                   // out: a barrier wait ran past its limit
};
size_t leaf_all_targs_bytes(int trees);
bool launch_leaf_all(int metric, bool screen, const float* x, int d, int ld, const _Float16* xh, int ldh,
                     const float* eps, const float* eps_max, const float* nh, const float* norms, const int32_t* perm,
                     const Leaf* items, const int64_t* item_off, int trees, int64_t n, int k, int32_t* ids,
                     float* vals, const int32_t* rs, const int32_t* rc, unsigned long long* rescored,
                     unsigned long long* screened, const LeafAllState& st, int device, hipStream_t s);

// ---- lists longer than 64 (big_k.hip) -----------------------------------------
// k in (64, kBigKMax]: lists in LDS, merged by sorting (list + a chunk of
// candidates) and keeping the first k distinct pairs.
constexpr int kBigKMax = 256;
// Every point's row merged with one tree's keys of its start range (the key
// table of launch_range_keys at offsets off[i]).
void launch_big_leaf_merge(const float* keys, const int64_t* off, const int32_t* perm, const int32_t* rs,
                           const int32_t* rc, int64_t n, int k, int32_t* ids, float* vals, hipStream_t s);
// One exact propagation round (EUCLIDEAN, TAXICAB, COSINE with carried
// distances: metrics whose stored value is the key).
void launch_big_propagate(int metric, const float* x, int d, int ld, const float* norms, int64_t n, int k,
                          const int32_t* ids_in, const float* vals_in, int32_t* ids_out, float* vals_out,
                          hipStream_t s);

// ---- FP64 exact k-NN and recall (the measurement reference) ------------------
// sq: FP64 squared norms (COSINE only, launch_sq64).  ids/dist: [nq][kq],
// sorted by (distance, id), the query excluded.
void launch_sq64(const float* x, int64_t n, int d, int ld, double* sq, hipStream_t s);
void launch_exact_knn(int metric, const float* x, int64_t n, int d, int ld, const double* sq, const int64_t* queries,
                      int64_t nq, int kq, int32_t* ids, double* dist, hipStream_t s);
// hits += graph neighbours (ids: [n][k]) of each query within its k-th exact distance kth[nq]
void launch_recall(int metric, const float* x, int d, int ld, const double* sq, const int64_t* queries, int64_t nq,
                   const double* kth, const int32_t* ids, int k, unsigned long long* hits, hipStream_t s);

// COSINE parity initial pass: ascending-id chunks of `cap` candidates over the
// union of all trees' start ranges, with the store-time flip after each chunk.
// Points are taken in the first tree's leaf order (perm); groups is a
// multiple of 8 and max_cand bounds every point's union.
struct CosineParityScratch {
  uint32_t* bits;     // [groups][words]
  uint32_t* summary;  // [groups][summary_words]
  int32_t* cand;      // [groups][max_cand]
  float* key;         // [groups][max_cand]
  float* keyarr;      // [groups][n] keys by candidate id (with a key table; else null)
  uint64_t* uo;       // [groups][8192] a point's distinct members in id order (with a key table; else null)
  int64_t words, summary_words, max_cand;
  int groups;
};
// ktab/koff (nullable): the per-tree key table of launch_range_keys and its
// offsets; without them the keys are computed from the rows per point.
void launch_cosine_parity_initial(const float* x, int d, int ld, const float* norms, int64_t n, int k, int trees,
                                  const int32_t* rs, const int32_t* rc, const int32_t* perm, int64_t cap,
                                  const CosineParityScratch& sc, int32_t* ids, float* vals, const float* ktab,
                                  const int64_t* koff, hipStream_t s);
// Dynamic LDS bytes of the COSINE parity pass: x_i (without the key table),
// the summary bitmap (n / 1024 bits) and the selection pool (k + cap pairs).
size_t cosine_parity_lds(int d, int64_t summary_words, int k, int64_t cap, bool table);
// One tree's keys of every owner of the work items against its own
// start-node range (any metric: the COSINE parity pass, and the leaf merge of
// lists longer than 64): key(i, perm[r]) at keys[off[i] + r - rs[i]];
// max_owners: the largest item size.
void launch_range_keys(int metric, const float* x, int d, int ld, const float* norms, const int32_t* perm,
                       const Leaf* items, int nitems, int max_owners, const int32_t* rs, const int32_t* rc,
                       const int64_t* off, float* keys, hipStream_t s);
// offsets[i] = counts[0] + ... + counts[i-1] (int64); block_sums needs
// ceil(n / 1024) entries; *total receives the sum.
void launch_count_scan(const int32_t* counts, int64_t n, int64_t* offsets, int64_t* block_sums, int64_t* total,
                       hipStream_t s);

}  // namespace aknn
