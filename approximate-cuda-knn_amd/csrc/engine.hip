// Host orchestration of the k-NN graph build on one GPU.
//
// calculate() is the reference's timed bracket (euclideandistancematrixgpu.cu:2861-2896):
// normalize, forest, stub + initial k-NN, propagation rounds.  Differences in
// structure (not in results):
//   * data stays row-major in HBM (one point = one contiguous row padded to a
//     multiple of 4 floats), so neighbour gathers are single 16-byte streams;
//   * the forest never leaves the device: no per-tree D2H/H2D round trip of
//     nodes and leaves (.cu:2653-2655, :1997-2065);
//   * the per-level host check of the reference (.cu:2602-2604) becomes one
//     check per batch of levels; levels past the end of a tree are no-ops and
//     the rand() stream is rewound to the first draw the tree did not use;
//   * one stream, no per-launch synchronisation.
#include "engine.hpp"

#include "directions.hpp"
#include "router.hpp"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <exception>
#include <thread>

namespace aknn {

namespace {
constexpr int kNumEvents = 8;
constexpr int kItemOwners = 64;  // owners per leaf k-NN work item (one wave)

// Work items of one tree's leaf k-NN from its leaves: {first owner position,
// owners (<= 64), candidate range start, count}.  A leaf with more than 64
// points gives one item per 64 owners over its own start range.  Consecutive
// small leaves are grouped while the item keeps at most 64 owners and the
// union of their start ranges stays within 1.25 times the largest one plus a
// tile (nested or shared start nodes: a cluster peeled off one point per
// level leaves a chain of one-point leaves whose ranges are nested); each
// pair is then masked by its owner's own range.
void build_work_items(std::vector<Leaf>& leaves, std::vector<Leaf>& items) {
  std::sort(leaves.begin(), leaves.end(), [](const Leaf& a, const Leaf& b) { return a.start < b.start; });
  items.clear();
  Leaf cur{0, 0, 0, 0};
  int32_t cur_max = 0;
  bool open = false;
  for (const Leaf& l : leaves) {
    if (l.size > kItemOwners) {
      if (open) items.push_back(cur);
      open = false;
      for (int32_t o = 0; o < l.size; o += kItemOwners)
        items.push_back(Leaf{l.start + o, std::min<int32_t>(kItemOwners, l.size - o), l.rstart, l.rcount});
      continue;
    }
    if (open && cur.size + l.size <= kItemOwners) {
      const int64_t u0 = std::min(cur.rstart, l.rstart);
      const int64_t u1 = std::max<int64_t>(static_cast<int64_t>(cur.rstart) + cur.rcount,
                                           static_cast<int64_t>(l.rstart) + l.rcount);
      const int32_t mx = std::max(cur_max, l.rcount);
      if (u1 - u0 <= static_cast<int64_t>(mx) + mx / 4 + 64) {
        cur.size += l.size;
        cur.rstart = static_cast<int32_t>(u0);
        cur.rcount = static_cast<int32_t>(u1 - u0);
        cur_max = mx;
        continue;
      }
    }
    if (open) items.push_back(cur);
    cur = l;
    cur_max = l.rcount;
    open = true;
  }
  if (open) items.push_back(cur);
}
}  // namespace

// Shapes a stage cannot run, refused before any work: a propagation round
// keeps one wave's hash set, candidate list and x_i in LDS; the COSINE parity
// pass keeps x_i in 64 KB of LDS.
static void check_shapes(const Params& p, int64_t n, int d) {
  if (p.k > 64) {
    // lists longer than 64 take the sort-merge kernels (big_k.hip), which
    // need the stored value to be the key
    if (p.metric == kCosine && !p.cosine_fixed)
      throw std::invalid_argument("k > 64 with COSINE needs option cosine_fixed (carried distances): the reference's "
                                  "flip of stored values is supported for k <= 64");
    return;
  }
  if (p.props > 0 && !propagation_fits(p.k, d))
    throw std::invalid_argument("propagation: k = " + std::to_string(p.k) + " with D = " + std::to_string(d) +
                                " needs more than 160 KB of LDS for one wave ((2(k^2+k+1) + k^2 + D) * 4 bytes)");
  if (p.metric == kCosine && !p.cosine_fixed) {
    // x_i, the summary bitmap and the selection pool in LDS
    const int64_t words = ((n + 31) / 32 + 31) / 32 * 32;
    const int64_t cap = std::min<int64_t>(static_cast<int64_t>(p.k) * p.k + p.k, n) - p.k;
    if (cosine_parity_lds(d, words / 32, p.k, cap, false) > 160 * 1024 - 4096)
      throw std::invalid_argument("COSINE (parity mode): D, N or k too large for LDS (x_i 4D, N/256 and 8(k^2+k) bytes)");
  }
}

Engine::Engine(int device) : device_(device) {
  int count = 0;
  AKNN_HIP_TRY(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) throw HipError("device " + std::to_string(device) + " is not present");
  AKNN_HIP_TRY(hipSetDevice(device_));
  AKNN_HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  for (int i = 0; i < kNumEvents; ++i) {
    hipEvent_t e;
    AKNN_HIP_TRY(hipEventCreate(&e));
    events_.push_back(e);
  }
  AKNN_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&ctl_host_), sizeof(ForestCtl), hipHostMallocDefault));
  AKNN_HIP_TRY(hipEventCreate(&pca_ev_));
  AKNN_HIP_TRY(hipEventCreate(&router_ev_));
  AKNN_HIP_TRY(hipStreamCreateWithFlags(&aux_, hipStreamNonBlocking));
}

Engine::~Engine() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  drop_round_graph();
  release();
  for (auto e : events_) (void)hipEventDestroy(e);
  for (auto e : launch_events_) (void)hipEventDestroy(e);
  if (pca_ev_) (void)hipEventDestroy(pca_ev_);
  if (router_ev_) (void)hipEventDestroy(router_ev_);
  if (aux_) {
    (void)hipStreamSynchronize(aux_);
    (void)hipStreamDestroy(aux_);
  }
  if (ctl_host_) (void)hipHostFree(ctl_host_);
  for (Pinned& b : pinned_)
    if (b.p) (void)hipHostFree(b.p);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void* Engine::pinned(int slot, size_t bytes) {
  Pinned& b = pinned_[slot];
  if (bytes > b.cap) {
    if (b.p) {
      AKNN_HIP_TRY(hipStreamSynchronize(stream_));  // no copy may still target the old buffer
      AKNN_HIP_TRY(hipHostFree(b.p));
      b.p = nullptr;
      b.cap = 0;
    }
    AKNN_HIP_TRY(hipHostMalloc(&b.p, bytes, hipHostMallocDefault));
    b.cap = bytes;
  }
  return b.p;
}

hipEvent_t Engine::ev(int i) { return events_.at(i); }

hipEvent_t Engine::launch_ev(int64_t i) {
  while (static_cast<int64_t>(launch_events_.size()) <= i) {
    hipEvent_t e;
    AKNN_HIP_TRY(hipEventCreate(&e));
    launch_events_.push_back(e);
  }
  return launch_events_[i];
}

template <typename T>
T* Engine::alloc(size_t count) {
  void* p = nullptr;
  AKNN_HIP_TRY(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
  allocs_.push_back(p);
  return static_cast<T*>(p);
}

void Engine::drop_round_graph() {
  if (round_exec_) (void)hipGraphExecDestroy(round_exec_);
  round_exec_ = nullptr;
  round_graph_mode_ = -1;
}

void Engine::release() {
  drop_round_graph();
  for (void* p : allocs_) (void)hipFree(p);
  allocs_.clear();
  x_ = x_raw_ = norms_ = mean_ = eps_ = nh_ = s8_ = e8_ = nullptr;
  mean64_ = nullptr;
  q8_ = nullptr;
  a8_ = nullptr;
  se8_ = nullptr;
  leaf_rescored_ = nullptr;
  xh_ = nullptr;
  gmax_ = nullptr;
  gids_[0] = gids_[1] = nullptr;
  gvals_[0] = gvals_[1] = nullptr;
  perm_ = nullptr;
  leaves_ = nullptr;
  rs_ = rc_ = nullptr;
  items_ = nullptr;
  items_cap_ = 0;
  leaf_st_ = LeafAllState{};
  leaf_all_used_ = false;
  ncand_ = nullptr;
  changed_[0] = changed_[1] = nullptr;
  fb_ = ForestBuffers{};
  cps_ = CosineParityScratch{};
  koff_ = kbsum_ = nullptr;
  ktab_ = nullptr;
  ktab_cap_ = 0;
  dirs_ = proj_ = nullptr;
  fbm_ = ForestBuffers{};
  fbm_trees_ = 0;
  proj_nd_ = 0;
  tree_ctl_ = nullptr;
  pca_p_ = pca_mom_ = pca_proj_ = pca_e_ = nullptr;
  pca_ = PcaBasis{};
  nd_ = 0;
  dirs_key_ = -1;
  bins_ = nullptr;
  bins_cap_ = 0;
  bin_bits_ = 0;
  bins_set_ = false;
}

void Engine::set_bins(const int32_t* bins, int bin_bits) {
  if (n_ <= 0) throw std::invalid_argument("set the points first");
  if (bin_bits < 1 || bin_bits > 20) throw std::invalid_argument("bin_bits must be in [1, 20]");
  for (int64_t i = 0; i < n_; ++i)
    if (bins[i] < 0 || bins[i] >= (1 << bin_bits)) throw std::invalid_argument("bin out of range");
  AKNN_HIP_TRY(hipSetDevice(device_));
  if (!bins_) {
    bins_ = alloc<int32_t>(n_);
    bins_cap_ = 1;
  }
  AKNN_HIP_TRY(hipMemcpy(bins_, bins, sizeof(int32_t) * n_, hipMemcpyHostToDevice));
  bin_bits_ = bin_bits;
  bins_set_ = true;
}

void Engine::set_bins_device(const int32_t* bins, int bin_bits) {
  if (bin_bits < 1 || bin_bits > 20) throw std::invalid_argument("bin_bits must be in [1, 20]");
  if (!bins_) {
    bins_ = alloc<int32_t>(n_);
    bins_cap_ = 1;
  }
  if (bins != bins_)
    AKNN_HIP_TRY(hipMemcpyAsync(bins_, bins, sizeof(int32_t) * n_, hipMemcpyDeviceToDevice, stream_));
  bin_bits_ = bin_bits;
  bins_set_ = true;
}

void Engine::set_data(const float* x_rowmajor, int64_t n, int d, const Params& p) {
  if (n < 1 || d < 1) throw std::invalid_argument("N and D must be positive");
  if (p.k < 1 || p.k > kBigKMax) throw std::invalid_argument("k must be in [1, " + std::to_string(kBigKMax) + "]");
  if (n < static_cast<int64_t>(p.k) + 1)
    throw std::invalid_argument("N must be at least k+1 (the reference walks past the root otherwise)");
  if (n > 0x7fffffffLL) throw std::invalid_argument("N must fit in int32");
  if (p.trees < 1) throw std::invalid_argument("the number of trees must be at least 1");
  if (p.leaf < 0 || p.props < 0) throw std::invalid_argument("leaf size and propagations must be >= 0");
  if (p.metric < kEuclidean || p.metric > kCosine) throw std::invalid_argument("unknown distance");
  if (p.split < 0 || p.split > 1) throw std::invalid_argument("split must be 0 (sign) or 1 (median)");
  check_shapes(p, n, d);
  AKNN_HIP_TRY(hipSetDevice(device_));
  AKNN_HIP_TRY(hipStreamSynchronize(stream_));
  release();
  p_ = p;
  prop_mode_ = -1;
  leaf_form_ = -1;
  n_ = n;
  d_ = d;
  ld_ = (d + 3) / 4 * 4;
  ldh_ = (d + 7) / 8 * 8;
  ld8_ = (d + 15) / 16 * 16;
  leaf_eff_ = std::max(p.leaf, 1);

  x_ = alloc<float>(static_cast<size_t>(n) * ld_);
  x_raw_ = alloc<float>(static_cast<size_t>(n) * ld_);
  AKNN_HIP_TRY(hipMemcpy2DAsync(x_raw_, ld_ * sizeof(float), x_rowmajor, d * sizeof(float), d * sizeof(float), n,
                                hipMemcpyHostToDevice, stream_));
  if (ld_ != d)
    AKNN_HIP_TRY(hipMemset2DAsync(reinterpret_cast<char*>(x_raw_) + d * sizeof(float), ld_ * sizeof(float), 0,
                                  (ld_ - d) * sizeof(float), n, stream_));
  if (p.metric == kCosine) norms_ = alloc<float>(n);
  mean_ = alloc<float>(d);
  gmax_ = alloc<uint32_t>(1);
  const size_t nk = static_cast<size_t>(n) * p.k;
  for (int b = 0; b < 2; ++b) {
    gids_[b] = alloc<int32_t>(nk);
    gvals_[b] = alloc<float>(nk);
  }
  perm_ = alloc<int32_t>(static_cast<size_t>(n) * p.trees);
  leaves_ = alloc<Leaf>(static_cast<size_t>(n) * p.trees);
  const int64_t blocks = (n + 1023) / 1024;
  fb_.n = n;
  fb_.max_buckets = n / (leaf_eff_ + 1) + 2;
  for (int b = 0; b < 2; ++b) {
    fb_.act_ids[b] = alloc<int32_t>(n);
    fb_.act_bkt[b] = alloc<int32_t>(n);
    fb_.bkt[b] = alloc<Bucket>(fb_.max_buckets);
  }
  fb_.rmask = alloc<uint64_t>(blocks * 16);
  fb_.wpre = alloc<int32_t>(blocks * 16);
  fb_.blk_r = alloc<int32_t>(blocks);
  fb_.blk_roff = alloc<int32_t>(blocks);
  fb_.bchild = alloc<int32_t>(fb_.max_buckets);
  fb_.piv[0] = alloc<int32_t>(2 * fb_.max_buckets);
  fb_.piv[1] = alloc<int32_t>(2 * fb_.max_buckets);
  fb_.bact = alloc<int32_t>(fb_.max_buckets);
  fb_.ctl = alloc<ForestCtl>(1);
  ncand_ = alloc<unsigned long long>(static_cast<size_t>(kCountWords) * std::max(p.props, 1));
  changed_[0] = alloc<uint8_t>(n);
  changed_[1] = alloc<uint8_t>(n);
  AKNN_HIP_TRY(hipStreamSynchronize(stream_));
}

void Engine::build_forest() {
  nleaves_.assign(p_.trees, 0);
  iters_.assign(p_.trees, 0);
  active_sum_ = 0;
  stats_forest = KernelStats{};
  if (fb_.amb_count) AKNN_HIP_TRY(hipMemsetAsync(fb_.amb_count, 0, sizeof(unsigned long long) * kSlotWords, stream_));
  const int n32 = static_cast<int>(n_);
  for (int t = 0; t < p_.trees; ++t) {
    const size_t base = draws_.pos();
    // root pivots (.cu:2481-2487)
    int32_t q1 = pick_index(draws_.at(base), n32);
    int32_t q2 = pick_index(draws_.at(base + 1), n32);
    if (q1 == q2) q2 = (q1 + 1) % n32;
    int32_t* perm = perm_ + static_cast<int64_t>(t) * n_;
    Leaf* leaves = leaves_ + static_cast<int64_t>(t) * n_;
    launch_forest_init(fb_, perm, leaves, n_, q1, q2, leaf_eff_, p_.k + 1, stream_);
    int level = 0;
    for (;;) {
      for (int b = 0; b < p_.level_batch; ++b, ++level) {
        const int32_t first = draws_.at(base + 2 + 2 * static_cast<size_t>(level));
        const int32_t second = draws_.at(base + 3 + 2 * static_cast<size_t>(level));
        // Leaf7(rand(), rand()) (.cu:2589): argument evaluation order is the host compiler's
        const int32_t r1 = p_.rng_order == 0 ? second : first;
        const int32_t r2 = p_.rng_order == 0 ? first : second;
        const int32_t v1 = static_cast<int32_t>((static_cast<double>(r1) / 2147483647.0) * 2147483647.0);
        const int32_t v2 = static_cast<int32_t>((static_cast<double>(r2) / 2147483647.0) * 2147483647.0);
        launch_forest_level(fb_, x_, d_, ld_, perm, leaves, level, v1, v2, leaf_eff_, p_.k + 1, stream_);
      }
      AKNN_HIP_TRY(hipMemcpyAsync(ctl_host_, fb_.ctl, sizeof(ForestCtl), hipMemcpyDeviceToHost, stream_));
      AKNN_HIP_TRY(hipStreamSynchronize(stream_));
      if (ctl_host_->done) break;
      if (level > n_ + 2) throw HipError("forest level loop did not terminate");
    }
    iters_[t] = ctl_host_->iterations;
    nleaves_[t] = ctl_host_->nleaves;
    active_sum_ += ctl_host_->active_sum;
    draws_.advance(2 + 2 * static_cast<size_t>(ctl_host_->iterations));
  }
  // algorithmic bytes: every active point's row (or its sketch row, radius
  // and norm, plus the exact row of each undecided point) and its bookkeeping
  // per level, plus the leaf array write (T * N * 8 with the leaf list)
  stats_forest.units = static_cast<double>(active_sum_) * (4.0 * d_ + 16.0) + 8.0 * n_ * p_.trees;
  if (fb_.amb_count) {
    unsigned long long amb = 0;
    slot_sums(fb_.amb_count, 1, &amb);
    stats_forest.examined = static_cast<double>(active_sum_);
    stats_forest.rescored = static_cast<double>(amb);
    // per active point its sketch row and 16 bytes of ids, bucket and output;
    // the radius and norm bounds are one value each per block
    stats_forest.units = static_cast<double>(active_sum_) * (2.0 * ldh_ + 16.0) + 4.0 * d_ * amb +
                         8.0 * n_ * p_.trees;
  }
}

// Median-split forest: per tree, the projections on its seeded directions
// (one pass over the points), then the level loop with the bin / median /
// position rules; no rand() draws.  The leaf list, start ranges and
// partition machinery are the sign-split forest's.
void Engine::build_forest_median() {
  nleaves_.assign(p_.trees, 0);
  iters_.assign(p_.trees, 0);
  active_sum_ = 0;
  stats_forest = KernelStats{};
  proj_dirs_ = 0;
  proj_launches_ = 0;
  const int nd = median_depth(n_, leaf_eff_);
  const int64_t key = (static_cast<int64_t>(p_.seed) << 20) ^ (static_cast<int64_t>(p_.trees) << 8) ^ nd;
  if (!dirs_ || nd != nd_ || key != dirs_key_) {
    if (!dirs_ || nd > nd_) dirs_ = alloc<float>(static_cast<size_t>(p_.trees) * nd * ld_);
    nd_ = nd;
    std::vector<float> all(static_cast<size_t>(p_.trees) * nd * ld_), one;
    for (int t = 0; t < p_.trees; ++t) {
      tree_directions(p_.seed, t, nd, d_, ld_, one);
      std::copy(one.begin(), one.end(), all.begin() + static_cast<size_t>(t) * nd * ld_);
    }
    AKNN_HIP_TRY(hipMemcpy(dirs_, all.data(), sizeof(float) * all.size(), hipMemcpyHostToDevice));
    dirs_key_ = key;
  }
  int nbins_sets = bins_set_ ? 1 : 0;
  if (routers_launched_) {
    // the learned leaf assignments of this data set, one bin array per
    // router, made on the side stream since the build began
    AKNN_HIP_TRY(hipStreamWaitEvent(stream_, router_ev_, 0));
    routers_launched_ = false;
    bin_bits_ = routers_[0]->bin_bits();
    bins_set_ = true;
    nbins_sets = static_cast<int>(routers_.size());
  }
  const int routed_trees = bins_set_ ? std::min(p_.router_trees, p_.trees) : 0;
  // Trees are independent here (no shared draw stream), so a batch of them is
  // built level by level together: one active list, one bucket list, three
  // passes per level for all of them, instead of per tree.
  const int bt = median_batch_trees(nd);
  MedianRules rules;
  rules.routed_trees = routed_trees;
  rules.bin_bits = bin_bits_;
  rules.nsets = std::max(nbins_sets, 1);
  rules.bins = bins_;
  rules.proj = proj_;
  rules.nd = nd;
  rules.n = n_;
  std::vector<int32_t> tl_host(2 * static_cast<size_t>(bt));
  for (int t0 = 0; t0 < p_.trees; t0 += bt) {
    const int nt = std::min(bt, p_.trees - t0);
    rules.tree0 = t0;
    int32_t* perm = perm_ + static_cast<int64_t>(t0) * n_;
    Leaf* leaves = leaves_ + static_cast<int64_t>(t0) * n_;
    // projections are made 16 directions at a time, the first time a level
    // needs one: routed trees use only the directions below their depth past
    // the bin levels.  A tree's first batch runs over all rows in order; a
    // later one only over the points still active then (the few of deep
    // buckets, of every tree of the batch).
    std::vector<int> projected(nt, 0);
    launch_forest_init_batch(fbm_, nt, perm, leaves, tree_ctl_, tree_ctl_ + bt, n_, leaf_eff_, stream_);
    int level = 0;
    for (;;) {
      for (int b = 0; b < p_.level_batch; ++b, ++level) {
        bool any_median = false;
        // trees whose first directions this level needs share passes over
        // all rows (up to kProjSetMax trees a pass)
        ProjSet set{};
        int groups = 0;
        auto flush = [&]() {
          if (groups == 0) return;
          launch_project_set(x_, n_, d_, ld_, set, groups, stream_);
          ++proj_launches_;
          groups = 0;
        };
        for (int tl = 0; tl < nt; ++tl) {
          const int bb = t0 + tl < routed_trees ? bin_bits_ : 0;
          const int dir = level - bb;
          if (dir < 0 || dir >= nd) continue;
          any_median = true;
          if (dir < projected[tl]) continue;
          const int cnt = std::min(kProjectStep, nd - projected[tl]);
          const float* dirs = dirs_ + (static_cast<size_t>(t0 + tl) * nd + projected[tl]) * ld_;
          float* proj = proj_ + (static_cast<int64_t>(tl) * nd + projected[tl]) * n_;
          if (projected[tl] == 0) {
            set.dirs[groups] = dirs;
            set.proj[groups] = proj;
            set.nd[groups] = cnt;
            if (++groups == kProjSetMax) flush();
          } else {
            launch_project(x_, n_, d_, ld_, dirs, cnt, proj, &fbm_, level, stream_);
            ++proj_launches_;
          }
          projected[tl] += cnt;
          proj_dirs_ += cnt;
        }
        flush();
        launch_forest_level_batch(fbm_, rules, any_median, level, perm, leaves, tree_ctl_, tree_ctl_ + bt, leaf_eff_,
                                  p_.k + 1, stream_);
      }
      AKNN_HIP_TRY(hipMemcpyAsync(ctl_host_, fbm_.ctl, sizeof(ForestCtl), hipMemcpyDeviceToHost, stream_));
      AKNN_HIP_TRY(hipStreamSynchronize(stream_));
      if (ctl_host_->done) break;
      if (level > n_ + nd + 32) throw HipError("forest level loop did not terminate");
    }
    AKNN_HIP_TRY(hipMemcpyAsync(tl_host.data(), tree_ctl_, sizeof(int32_t) * 2 * bt, hipMemcpyDeviceToHost, stream_));
    AKNN_HIP_TRY(hipStreamSynchronize(stream_));
    for (int tl = 0; tl < nt; ++tl) {
      nleaves_[t0 + tl] = tl_host[tl];
      // levels that split a bucket, plus the last one that found none
      iters_[t0 + tl] = tl_host[bt + tl] + 2;
    }
    active_sum_ += ctl_host_->active_sum;
  }
  // algorithmic bytes: one pass over the rows and the projection writes per
  // tree, then per level and active point its projection, id and bucket
  // bookkeeping, plus the leaf array write
  stats_forest.units = static_cast<double>(proj_launches_) * n_ * 4.0 * d_ +
                       static_cast<double>(proj_dirs_) * n_ * 4.0 + static_cast<double>(active_sum_) * 20.0 +
                       8.0 * n_ * p_.trees;
}

// Trees of a median-split batch: as many as the batch buffers (16 bytes of
// active lists and 4 * nd of projections per point and tree, plus buckets)
// can hold in a quarter of the free device memory, at most 64 and at most
// 2^31 - 1 active points.  Allocated on the first build of a data set.
int Engine::median_batch_trees(int nd) {
  if (fbm_.ctl != nullptr && nd <= proj_nd_) return fbm_trees_;
  const int64_t per_tree = n_ * (16 + 4 * static_cast<int64_t>(nd) + 1) + 128 * (n_ / (leaf_eff_ + 1) + 2);
  size_t free_b = 0, total_b = 0;
  AKNN_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
  int64_t bt = std::min<int64_t>({p_.trees, 64, 0x7fffffffLL / n_,
                                  static_cast<int64_t>(free_b / 4) / std::max<int64_t>(per_tree, 1)});
  bt = std::max<int64_t>(bt, 1);
  const int64_t cap = bt * n_;
  const int64_t blocks = (cap + 1023) / 1024;
  fbm_ = ForestBuffers{};
  fbm_.n = cap;
  fbm_.max_buckets = bt * (n_ / (leaf_eff_ + 1) + 2);
  for (int b = 0; b < 2; ++b) {
    fbm_.act_ids[b] = alloc<int32_t>(cap);
    fbm_.act_bkt[b] = alloc<int32_t>(cap);
    fbm_.bkt[b] = alloc<Bucket>(fbm_.max_buckets);
  }
  fbm_.rmask = alloc<uint64_t>(blocks * 16);
  fbm_.wpre = alloc<int32_t>(blocks * 16);
  fbm_.blk_r = alloc<int32_t>(blocks);
  fbm_.blk_roff = alloc<int32_t>(blocks);
  fbm_.bchild = alloc<int32_t>(fbm_.max_buckets);
  fbm_.bact = alloc<int32_t>(fbm_.max_buckets);
  fbm_.ctl = alloc<ForestCtl>(1);
  tree_ctl_ = alloc<int32_t>(2 * bt);
  proj_ = alloc<float>(static_cast<size_t>(bt) * nd * n_);
  proj_nd_ = nd;
  fbm_trees_ = static_cast<int>(bt);
  return fbm_trees_;
}

void Engine::initial_knn() {
  stats_leaf_knn = KernelStats{};
  stats_leaf_knn.segments = p_.trees;
  tree_screened_.assign(p_.trees, 0);
  leaf_all_used_ = false;
  int32_t* ids = gids_[cur_];
  float* vals = gvals_[cur_];
  launch_stub_init(ids, vals, n_ * p_.k, stream_);
  const bool parity_cosine = p_.metric == kCosine && !p_.cosine_fixed;
  if (parity_cosine) {
    if (!rs_) {
      rs_ = alloc<int32_t>(static_cast<size_t>(n_) * p_.trees);
      rc_ = alloc<int32_t>(static_cast<size_t>(n_) * p_.trees);
    }
    for (int t = 0; t < p_.trees; ++t)
      launch_leaf_ranges(perm_ + t * n_, leaves_ + t * n_, nleaves_[t], rs_ + t * n_, rc_ + t * n_, stream_);
    // the leaf lists of all trees in one round trip
    std::vector<int64_t> leaf_off(p_.trees + 1, 0);
    for (int t = 0; t < p_.trees; ++t) leaf_off[t + 1] = leaf_off[t] + nleaves_[t];
    Leaf* lh = static_cast<Leaf*>(pinned(0, sizeof(Leaf) * static_cast<size_t>(std::max<int64_t>(leaf_off[p_.trees], 1))));
    for (int t = 0; t < p_.trees; ++t)
      AKNN_HIP_TRY(hipMemcpyAsync(lh + leaf_off[t], leaves_ + t * n_, sizeof(Leaf) * nleaves_[t], hipMemcpyDeviceToHost,
                                  stream_));
    // offsets of every (tree, point) range in the key table: the exclusive
    // prefix sum of the range sizes, tree after tree
    const int64_t tn = static_cast<int64_t>(n_) * p_.trees;
    if (!koff_) {
      koff_ = alloc<int64_t>(tn + 1);
      kbsum_ = alloc<int64_t>((tn + 1023) / 1024 + 1);
    }
    launch_count_scan(rc_, tn, koff_, kbsum_, koff_ + tn, stream_);
    int64_t ktotal = 0;
    AKNN_HIP_TRY(hipMemcpyAsync(&ktotal, koff_ + tn, sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
    AKNN_HIP_TRY(hipStreamSynchronize(stream_));
    // a point's union holds at most the sum over trees of the largest start
    // range: that bounds each group's candidate scratch
    int64_t bound = 0;
    for (int t = 0; t < p_.trees; ++t) {
      int32_t mx = 0;
      for (int64_t l = leaf_off[t]; l < leaf_off[t + 1]; ++l) mx = std::max(mx, lh[l].rcount);
      bound += mx;
    }
    const int64_t need = std::min<int64_t>(n_, std::max<int64_t>(bound, 1));
    // The key table (4 bytes per range member) and each group's key-by-id
    // array (4n bytes) are used when they fit in 70% of the free device
    // memory.  Each group works through its points one after another, so
    // 2048 of them (8 per CU) keep the device busy: their scratch takes up to
    // 24 GB with the table, 4 GB without.
    // bitmap words per group padded to whole 32-word lines (the enumeration
    // reads a line per summary bit); summary bits live in LDS
    const int64_t words = ((n_ + 31) / 32 + 31) / 32 * 32, summary = words / 32;
    auto groups_for = [&](bool table) {
      const int64_t per_group = 4 * (words + summary) + 8 * need + (table ? 4 * n_ + 8 * 8192 : 0);
      const int64_t fit = (int64_t{table ? 24 : 4} << 30) / std::max<int64_t>(per_group, 1);
      return static_cast<int>(std::max<int64_t>(8, std::min<int64_t>({2048, fit / 8 * 8, (n_ + 7) / 8 * 8})));
    };
    bool table = p_.cosine_table != 0;
    if (table) {
      size_t free_b = 0, total_b = 0;
      AKNN_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
      const int64_t have_tab = ktab_cap_ >= ktotal ? 0 : ktotal * 4;
      const int64_t have_arr = cps_.keyarr && cps_.max_cand >= need ? 0 : int64_t{groups_for(true)} * n_ * 4;
      table = static_cast<double>(have_tab + have_arr) <= 0.7 * static_cast<double>(free_b);
    }
    if (cps_.bits && (need > cps_.max_cand || (table && !cps_.keyarr))) cps_ = CosineParityScratch{};  // grow (old buffers freed at release)
    if (!cps_.bits) {
      cps_.groups = groups_for(table);
      cps_.words = words;
      cps_.summary_words = summary;
      cps_.max_cand = need;
      cps_.bits = alloc<uint32_t>(static_cast<size_t>(cps_.groups) * cps_.words);
      cps_.summary = nullptr;
      cps_.cand = alloc<int32_t>(static_cast<size_t>(cps_.groups) * cps_.max_cand);
      cps_.key = alloc<float>(static_cast<size_t>(cps_.groups) * cps_.max_cand);
      cps_.keyarr = table ? alloc<float>(static_cast<size_t>(cps_.groups) * n_) : nullptr;
      cps_.uo = table ? alloc<uint64_t>(static_cast<size_t>(cps_.groups) * 8192) : nullptr;
      AKNN_HIP_TRY(hipMemsetAsync(cps_.bits, 0, sizeof(uint32_t) * cps_.groups * cps_.words, stream_));
    }
    table = table && cps_.keyarr != nullptr;
    if (table && ktab_cap_ < ktotal) {
      ktab_ = alloc<float>(static_cast<size_t>(ktotal));
      ktab_cap_ = ktotal;
    }
    const int64_t dim_idx = std::min<int64_t>(static_cast<int64_t>(p_.k) * p_.k + p_.k, n_);
    AKNN_HIP_TRY(hipEventRecord(launch_ev(0), stream_));
    if (table) {
      // every tree's keys, per work item (the leaf k-NN's items: big leaves
      // cut into 64-owner items, runs of small ones grouped)
      std::vector<int64_t> item_off(p_.trees + 1, 0);
      for (int t = 0; t < p_.trees; ++t) item_off[t + 1] = item_off[t] + nleaves_[t] + n_ / kItemOwners + 1;
      const size_t ni = static_cast<size_t>(item_off[p_.trees]);
      if (ni > items_cap_) {
        items_ = alloc<Leaf>(ni);
        items_cap_ = ni;
      }
      Leaf* ih = static_cast<Leaf*>(pinned(1, sizeof(Leaf) * ni));
      std::vector<Leaf> host, its;
      for (int t = 0; t < p_.trees; ++t) {
        host.assign(lh + leaf_off[t], lh + leaf_off[t + 1]);
        build_work_items(host, its);
        const int nit = static_cast<int>(its.size());
        if (nit > item_off[t + 1] - item_off[t]) throw HipError("leaf work items exceed their bound");
        std::copy(its.begin(), its.end(), ih + item_off[t]);
        AKNN_HIP_TRY(hipMemcpyAsync(items_ + item_off[t], ih + item_off[t], sizeof(Leaf) * nit, hipMemcpyHostToDevice,
                                    stream_));
        launch_range_keys(kCosine, x_, d_, ld_, norms_, perm_ + t * n_, items_ + item_off[t], nit, kItemOwners, rs_ + t * n_,
                           rc_ + t * n_, koff_ + t * n_, ktab_, stream_);
      }
    }
    launch_cosine_parity_initial(x_, d_, ld_, norms_, n_, p_.k, p_.trees, rs_, rc_, perm_, dim_idx - p_.k, cps_, ids,
                                 vals, table ? ktab_ : nullptr, table ? koff_ : nullptr, stream_);
    AKNN_HIP_TRY(hipEventRecord(launch_ev(1), stream_));
    stats_leaf_knn.launches = 1;
    cosine_table_used_ = table;
    return;
  }
  if (p_.k > 64) {
    initial_knn_big();
    return;
  }
  // Work items per tree (host, from the leaf lists): leaves in position order,
  // big ones cut into 64-owner items, runs of small ones whose start ranges
  // mostly overlap grouped into one item scanning the union of the ranges.
  if (!rs_) {
    rs_ = alloc<int32_t>(static_cast<size_t>(n_) * p_.trees);
    rc_ = alloc<int32_t>(static_cast<size_t>(n_) * p_.trees);
  }
  // The leaf lists of all trees come over in one round trip (asynchronous
  // copies into one pinned buffer, one sync).  Each tree's work items are
  // then built on the host while the device runs the previous tree, and go
  // up just before its kernel.  A leaf gives at most ceil(size / 64) items,
  // so every tree's items fit in nleaves + n / 64 + 1 entries.
  std::vector<std::vector<Leaf>>& items = items_host_;
  items.assign(p_.trees, {});
  std::vector<int64_t> item_off(p_.trees + 1, 0);
  for (int t = 0; t < p_.trees; ++t) item_off[t + 1] = item_off[t] + nleaves_[t] + n_ / kItemOwners + 1;
  Leaf* lh = nullptr;
  std::vector<int64_t> leaf_off(p_.trees + 1, 0);
  {
    for (int t = 0; t < p_.trees; ++t) leaf_off[t + 1] = leaf_off[t] + nleaves_[t];
    lh = static_cast<Leaf*>(pinned(0, sizeof(Leaf) * static_cast<size_t>(std::max<int64_t>(leaf_off[p_.trees], 1))));
    for (int t = 0; t < p_.trees; ++t)
      AKNN_HIP_TRY(hipMemcpyAsync(lh + leaf_off[t], leaves_ + t * n_, sizeof(Leaf) * nleaves_[t],
                                  hipMemcpyDeviceToHost, stream_));
    AKNN_HIP_TRY(hipStreamSynchronize(stream_));
    for (int t = 0; t < p_.trees; ++t)
      launch_leaf_ranges(perm_ + t * n_, leaves_ + t * n_, nleaves_[t], rs_ + t * n_, rc_ + t * n_, stream_);
    const size_t ni = static_cast<size_t>(item_off[p_.trees]);
    if (ni > items_cap_) {
      items_ = alloc<Leaf>(ni);
      items_cap_ = ni;
    }
  }
  Leaf* ih = static_cast<Leaf*>(pinned(1, sizeof(Leaf) * static_cast<size_t>(item_off[p_.trees])));
  bool mfma = leaf_mfma_active();
  if (mfma) {
    AKNN_HIP_TRY(hipMemsetAsync(leaf_rescored_, 0, sizeof(unsigned long long) * kSlotWords, stream_));
    AKNN_HIP_TRY(hipMemsetAsync(leaf_rescored_ + 2 * kSlotWords, 0, sizeof(unsigned long long) * kSlotWords, stream_));
  }
  // leaf_batch 2: the launch form learned on this data set.  Unknown: the
  // all-trees launch runs first (a refused one leaves the stub rows as they
  // are), then the same rows again from stubs one launch per tree, and the
  // faster form is kept.
  const int form = p_.leaf_batch == 2 ? leaf_form_ : p_.leaf_batch;
  if (form == 1 && initial_knn_all(lh, leaf_off, mfma)) return;
  bool tune = form < 0;
  float all_ms = 0;
  if (tune) {
    if (initial_knn_all(lh, leaf_off, mfma)) {
      // its item staging is reused below and its error flag read: all done
      AKNN_HIP_TRY(hipStreamSynchronize(stream_));
      leaf_all_stats(mfma);
      AKNN_HIP_TRY(hipEventElapsedTime(&all_ms, ev(7), launch_ev(1)));
      leaf_all_used_ = false;
      stats_leaf_knn = KernelStats{};
      stats_leaf_knn.segments = p_.trees;
      tree_screened_.assign(p_.trees, 0);
      launch_stub_init(ids, vals, n_ * p_.k, stream_);
      if (mfma) {
        AKNN_HIP_TRY(hipMemsetAsync(leaf_rescored_, 0, sizeof(unsigned long long) * kSlotWords, stream_));
        AKNN_HIP_TRY(hipMemsetAsync(leaf_rescored_ + 2 * kSlotWords, 0, sizeof(unsigned long long) * kSlotWords,
                                    stream_));
      }
      AKNN_HIP_TRY(hipEventRecord(ev(4), stream_));
    } else {
      tune = false;
      leaf_form_ = 0;
    }
  }
  // leaf-screen counters (re-scored, -, screened slots) copied after each tree
  const size_t cw = 3 * static_cast<size_t>(kSlotWords);
  unsigned long long* ch = static_cast<unsigned long long*>(pinned(2, sizeof(unsigned long long) * cw * p_.trees));
  auto counts = [&](int t, unsigned long long c[2]) {
    c[0] = c[1] = 0;
    for (int s = 0; s < kCountSlots; ++s) {
      c[0] += ch[cw * t + s * kCountStride];
      c[1] += ch[cw * t + 2 * kSlotWords + s * kCountStride];
    }
  };
  for (int t = 0; t < p_.trees; ++t) {
    // The screen pays when few pairs survive it: after a screened tree that
    // re-scored more than 30% of its pairs, the remaining trees take the
    // plain kernel (same graph).  Tree t - 2's counts decide for tree t: they
    // were copied behind its kernel, so waiting for them does not let the
    // device run dry (tree t - 1 is queued behind them).
    if (mfma && t > 2) {
      AKNN_HIP_TRY(hipEventSynchronize(launch_ev(2 * (t - 1))));
      unsigned long long a[2], b[2];
      counts(t - 2, a);
      counts(t - 3, b);
      const unsigned long long dres = a[0] - b[0], dscr = a[1] - b[1];
      if (dscr > 0 && 10 * dres > 3 * dscr) mfma = false;
    }
    tree_screened_[t] = mfma && t > 0 ? 1 : 0;
    {
      std::vector<Leaf> host(lh + leaf_off[t], lh + leaf_off[t + 1]);
      build_work_items(host, items[t]);
    }
    const int nit = static_cast<int>(items[t].size());
    if (nit > item_off[t + 1] - item_off[t]) throw HipError("leaf work items exceed their bound");
    std::copy(items[t].begin(), items[t].end(), ih + item_off[t]);
    Leaf* it = items_ + item_off[t];
    AKNN_HIP_TRY(hipMemcpyAsync(it, ih + item_off[t], sizeof(Leaf) * nit, hipMemcpyHostToDevice, stream_));
    AKNN_HIP_TRY(hipEventRecord(launch_ev(2 * t), stream_));
    // the first tree meets stub rows (nothing to screen against): plain kernel
    if (mfma && t > 0)
      launch_leaf_mfma(kmetric(), x_, d_, ld_, xh_, ldh_, eps_, eps_ + n_, nh_, norms_, perm_ + t * n_, it, nit, kItemOwners,
                       p_.k, ids, vals, rs_ + t * n_, rc_ + t * n_, leaf_rescored_, leaf_rescored_ + 2 * kSlotWords,
                       stream_);
    else
      launch_leaf_knn(kmetric(), x_, d_, ld_, norms_, perm_ + t * n_, it, nit, kItemOwners, p_.k, ids, vals,
                      rs_ + t * n_, rc_ + t * n_, stream_);
    AKNN_HIP_TRY(hipEventRecord(launch_ev(2 * t + 1), stream_));
    if (mfma)
      AKNN_HIP_TRY(hipMemcpyAsync(ch + cw * t, leaf_rescored_, sizeof(unsigned long long) * cw,
                                  hipMemcpyDeviceToHost, stream_));
  }
  stats_leaf_knn.launches = p_.trees;
  if (tune) {
    AKNN_HIP_TRY(hipEventRecord(ev(5), stream_));
    AKNN_HIP_TRY(hipEventSynchronize(ev(5)));
    float per_tree = 0;
    AKNN_HIP_TRY(hipEventElapsedTime(&per_tree, ev(4), ev(5)));
    leaf_tune_ms_[0] = per_tree;
    leaf_tune_ms_[1] = all_ms;
    leaf_form_ = all_ms < per_tree ? 1 : 0;
  }
}

// Lists longer than 64 (big_k.hip), tree by tree: each tree's keys of every
// point against its start range into a table (the leaf kernels' tiling over
// the work items), then every point's row merged with its table row.
void Engine::initial_knn_big() {
  if (!rs_) {
    rs_ = alloc<int32_t>(static_cast<size_t>(n_) * p_.trees);
    rc_ = alloc<int32_t>(static_cast<size_t>(n_) * p_.trees);
  }
  const int64_t tn = static_cast<int64_t>(n_) * p_.trees;
  if (!koff_) {
    koff_ = alloc<int64_t>(tn + 1);
    kbsum_ = alloc<int64_t>((tn + 1023) / 1024 + 1);
  }
  std::vector<int64_t> leaf_off(p_.trees + 1, 0);
  for (int t = 0; t < p_.trees; ++t) leaf_off[t + 1] = leaf_off[t] + nleaves_[t];
  Leaf* lh = static_cast<Leaf*>(pinned(0, sizeof(Leaf) * static_cast<size_t>(std::max<int64_t>(leaf_off[p_.trees], 1))));
  for (int t = 0; t < p_.trees; ++t)
    AKNN_HIP_TRY(hipMemcpyAsync(lh + leaf_off[t], leaves_ + t * n_, sizeof(Leaf) * nleaves_[t], hipMemcpyDeviceToHost,
                                stream_));
  AKNN_HIP_TRY(hipStreamSynchronize(stream_));
  const size_t cap_items = static_cast<size_t>(*std::max_element(nleaves_.begin(), nleaves_.end())) + n_ / kItemOwners + 1;
  if (cap_items > items_cap_) {
    items_ = alloc<Leaf>(cap_items);
    items_cap_ = cap_items;
  }
  Leaf* ih = static_cast<Leaf*>(pinned(1, sizeof(Leaf) * cap_items));
  int32_t* ids = gids_[cur_];
  float* vals = gvals_[cur_];
  std::vector<Leaf> host, its;
  for (int t = 0; t < p_.trees; ++t) {
    int32_t* rs = rs_ + t * n_;
    int32_t* rc = rc_ + t * n_;
    launch_leaf_ranges(perm_ + t * n_, leaves_ + t * n_, nleaves_[t], rs, rc, stream_);
    launch_count_scan(rc, n_, koff_, kbsum_, koff_ + n_, stream_);
    int64_t total = 0;
    AKNN_HIP_TRY(hipMemcpyAsync(&total, koff_ + n_, sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
    host.assign(lh + leaf_off[t], lh + leaf_off[t + 1]);
    build_work_items(host, its);
    // the previous tree's items and table are in use until the stream gets here
    AKNN_HIP_TRY(hipStreamSynchronize(stream_));
    if (total > ktab_cap_) {
      ktab_ = alloc<float>(static_cast<size_t>(total));
      ktab_cap_ = total;
    }
    std::copy(its.begin(), its.end(), ih);
    AKNN_HIP_TRY(hipMemcpyAsync(items_, ih, sizeof(Leaf) * its.size(), hipMemcpyHostToDevice, stream_));
    AKNN_HIP_TRY(hipEventRecord(launch_ev(2 * t), stream_));
    launch_range_keys(kmetric(), x_, d_, ld_, norms_, perm_ + t * n_, items_, static_cast<int>(its.size()),
                      kItemOwners, rs, rc, koff_, ktab_, stream_);
    launch_big_leaf_merge(ktab_, koff_, perm_ + t * n_, rs, rc, n_, p_.k, ids, vals, stream_);
    AKNN_HIP_TRY(hipEventRecord(launch_ev(2 * t + 1), stream_));
  }
  stats_leaf_knn.launches = p_.trees;
}

// Propagation rounds for lists longer than 64 (big_k.hip): every round
// exact, every neighbour expanded.
void Engine::propagate_big() {
  const int base = 2 * p_.trees;
  round_screened_.assign(std::max(p_.props, 0), 0);
  round_form_.assign(std::max(p_.props, 0), 0);
  rounds_from_graph_ = false;
  AKNN_HIP_TRY(hipMemsetAsync(ncand_, 0, sizeof(unsigned long long) * kCountWords * std::max(p_.props, 1), stream_));
  for (int r = 0; r < p_.props; ++r) {
    AKNN_HIP_TRY(hipEventRecord(launch_ev(base + 2 * r), stream_));
    launch_big_propagate(kmetric(), x_, d_, ld_, norms_, n_, p_.k, gids_[cur_], gvals_[cur_], gids_[cur_ ^ 1],
                         gvals_[cur_ ^ 1], stream_);
    AKNN_HIP_TRY(hipEventRecord(launch_ev(base + 2 * r + 1), stream_));
    cur_ ^= 1;
  }
  stats_propagate.launches = p_.props;
}

// Every tree's leaf k-NN in one persistent launch (launch_leaf_all).  The
// work items of all trees are built first, on host threads in parallel, and
// go up in one copy with their offsets; the launch's tree modes, clock stamps
// and error flag come back behind it.  False when the device cannot hold the
// launch's grid (nothing launched).
bool Engine::initial_knn_all(const Leaf* lh, const std::vector<int64_t>& leaf_off, bool mfma) {
  const int T = p_.trees;
  // the tuning compares from here: the host's item building holds the device
  AKNN_HIP_TRY(hipEventRecord(ev(7), stream_));
  std::vector<std::vector<Leaf>>& items = items_host_;
  items.assign(T, {});
  {
    const int nth = std::min(T, 16);
    std::vector<std::thread> th;
    for (int w = 0; w < nth; ++w)
      th.emplace_back([&, w] {
        std::vector<Leaf> host;
        for (int t = w; t < T; t += nth) {
          host.assign(lh + leaf_off[t], lh + leaf_off[t + 1]);
          build_work_items(host, items[t]);
        }
      });
    for (std::thread& x : th) x.join();
  }
  std::vector<int64_t> off(T + 1, 0);
  for (int t = 0; t < T; ++t) {
    if (static_cast<int64_t>(items[t].size()) > nleaves_[t] + n_ / kItemOwners + 1)
      throw HipError("leaf work items exceed their bound");
    off[t + 1] = off[t] + static_cast<int64_t>(items[t].size());
  }
  if (static_cast<size_t>(off[T]) > items_cap_) throw HipError("leaf work items exceed their buffer");
  Leaf* ih = static_cast<Leaf*>(pinned(1, sizeof(Leaf) * static_cast<size_t>(std::max<int64_t>(off[T], 1))));
  for (int t = 0; t < T; ++t) std::copy(items[t].begin(), items[t].end(), ih + off[t]);
  // slot 3: modes [T] (padded to 8 bytes) | stamps [T + 1] | error
  const size_t b_mode = sizeof(int32_t) * T, b_st = sizeof(uint64_t) * (T + 1);
  const size_t b_mode8 = (b_mode + 7) / 8 * 8;
  char* h3 = static_cast<char*>(pinned(3, b_mode8 + b_st + sizeof(int)));
  if (!leaf_st_.queue) {
    leaf_st_.targs = alloc<char>(leaf_all_targs_bytes(T));
    leaf_st_.queue = alloc<unsigned>(8 * static_cast<size_t>(T));
    leaf_st_.bar = alloc<unsigned>(2);
    leaf_st_.mode = alloc<int32_t>(T);
    leaf_st_.stamps = alloc<unsigned long long>(T + 1);
    leaf_st_.prev = alloc<unsigned long long>(2);
    leaf_st_.error = alloc<int>(1);
    AKNN_HIP_TRY(hipDeviceGetAttribute(&wall_khz_, hipDeviceAttributeWallClockRate, device_));
  }
  AKNN_HIP_TRY(hipMemcpyAsync(items_, ih, sizeof(Leaf) * off[T], hipMemcpyHostToDevice, stream_));
  AKNN_HIP_TRY(hipEventRecord(launch_ev(0), stream_));
  const bool ok = launch_leaf_all(kmetric(), mfma, x_, d_, ld_, xh_, ldh_, eps_, eps_ ? eps_ + n_ : nullptr, nh_,
                                  norms_, perm_, items_, off.data(), T, n_, p_.k, gids_[cur_], gvals_[cur_], rs_,
                                  rc_, mfma ? leaf_rescored_ : nullptr, mfma ? leaf_rescored_ + 2 * kSlotWords : nullptr,
                                  leaf_st_, device_, stream_);
  if (!ok) return false;
  AKNN_HIP_TRY(hipEventRecord(launch_ev(1), stream_));
  char* hm = h3;
  char* hs = hm + b_mode8;
  AKNN_HIP_TRY(hipMemcpyAsync(hm, leaf_st_.mode, b_mode, hipMemcpyDeviceToHost, stream_));
  AKNN_HIP_TRY(hipMemcpyAsync(hs, leaf_st_.stamps, b_st, hipMemcpyDeviceToHost, stream_));
  AKNN_HIP_TRY(hipMemcpyAsync(hs + b_st, leaf_st_.error, sizeof(int), hipMemcpyDeviceToHost, stream_));
  stats_leaf_knn.launches = 1;
  leaf_all_used_ = true;
  return true;
}

// After the stream has finished: the all-trees launch's tree modes and
// per-tree times (device clock stamps at its barriers).
void Engine::leaf_all_stats(bool mfma) {
  const int T = p_.trees;
  const size_t b_mode8 = (sizeof(int32_t) * T + 7) / 8 * 8, b_st = sizeof(uint64_t) * (T + 1);
  const char* h3 = static_cast<const char*>(pinned_[3].p);
  const int32_t* mode = reinterpret_cast<const int32_t*>(h3);
  const char* hs = h3 + b_mode8;
  uint64_t st[2];
  int err = 0;
  std::memcpy(&err, hs + b_st, sizeof(int));
  if (err) throw HipError("leaf k-NN: a grid barrier wait ran past its limit");
  for (int t = 0; t < T; ++t) {
    std::memcpy(st, hs + sizeof(uint64_t) * t, sizeof(st));
    const double ms = static_cast<double>(st[1] - st[0]) / std::max(wall_khz_, 1);
    tree_screened_[t] = mfma && mode[t] ? 1 : 0;
    if (tree_screened_[t]) {
      stats_leaf_knn.screened_ms += ms;
      stats_leaf_knn.screened_launches += 1;
    }
  }
}

// Arguments of propagation round r, which reads graph buffer `cur` and
// writes the other one.
static PropagateArgs round_args(const Params& p, int kmetric, bool flip, bool screen, const PcaBasis& pb,
                                float* pca_proj, float* pca_e, float* x, int d, int ld,
                                _Float16* xh, int ldh, float* eps, int8_t* q8, int ld8, float* s8, int32_t* a8,
                                float* e8, float* norms, int32_t* order, int64_t n, int32_t* const gids[2],
                                float* const gvals[2], int cur, int r, unsigned long long* ncand,
                                uint8_t* const changed[2]) {
  PropagateArgs a;
  a.metric = kmetric;
  a.flip = flip;
  a.x = x;
  a.d = d;
  a.ld = ld;
  if (screen) {
    a.xh = xh;
    a.ldh = ldh;
    a.eps = eps;
    // the int8 sketch: EUCLIDEAN rounds with sketch_int8; COSINE parity
    // rounds (the flip) with flip_int8 whenever the int8 rows were made
    if ((p.sketch_int8 && p.metric == kEuclidean) || (flip && p.flip_int8 && q8 != nullptr)) {
      a.q8 = q8;
      a.ld8 = ld8;
      a.s8 = s8;
      a.a8 = a8;
      a.e8 = e8;
    }
  }
  if (pb.r > 0 && pca_proj != nullptr) {
    // the principal-subspace screen replaces the sketch screens
    a.xh = nullptr;
    a.eps = nullptr;
    a.q8 = nullptr;
    if (p.pca_fp16)
      a.pca16 = reinterpret_cast<const _Float16*>(pca_proj);  // the buffer holds FP16 rows then
    else
      a.pca = pca_proj;
    a.pca_r = pb.r;
    a.pca_e = pca_e;
    a.pca_sigma = pb.sigma;
  }
  a.norms = norms;
  a.order = order;
  a.n = n;
  a.k = p.k;
  a.ids_in = gids[cur];
  a.vals_in = gvals[cur];
  a.ids_out = gids[cur ^ 1];
  a.vals_out = gvals[cur ^ 1];
  a.ncand = ncand + static_cast<size_t>(kCountWords) * r;
  if (p.skip_settled) {
    a.hist.ids_prev = r > 0 ? gids[cur ^ 1] : nullptr;
    a.hist.changed_prev = changed[(r + 1) & 1];
    a.hist.changed_out = changed[r & 1];
  }
  return a;
}

// Totals of `kinds` consecutive spread counters (kSlotWords u64 each) on the
// device (synchronises the stream).
void Engine::slot_sums(const unsigned long long* dev, int kinds, unsigned long long* out) {
  std::vector<unsigned long long> w(static_cast<size_t>(kSlotWords) * kinds);
  AKNN_HIP_TRY(hipMemcpyAsync(w.data(), dev, sizeof(unsigned long long) * w.size(), hipMemcpyDeviceToHost, stream_));
  AKNN_HIP_TRY(hipStreamSynchronize(stream_));
  for (int k = 0; k < kinds; ++k) {
    out[k] = 0;
    for (int s = 0; s < kCountSlots; ++s) out[k] += w[static_cast<size_t>(k) * kSlotWords + s * kCountStride];
  }
}

// Sums of round r's counter slots: out[0] examined, out[1] re-scored (synchronises).
void Engine::round_counts(int r, unsigned long long out[2]) {
  std::vector<unsigned long long> w(kCountWords);
  AKNN_HIP_TRY(hipMemcpyAsync(w.data(), ncand_ + static_cast<size_t>(kCountWords) * r, sizeof(unsigned long long) * w.size(),
                              hipMemcpyDeviceToHost, stream_));
  AKNN_HIP_TRY(hipStreamSynchronize(stream_));
  out[0] = out[1] = 0;
  for (int s = 0; s < kCountSlots; ++s) {
    out[0] += w[static_cast<size_t>(s) * kCountStride];
    out[1] += w[static_cast<size_t>(kCountSlots + s) * kCountStride];
  }
}

void Engine::propagate() {
  stats_propagate = KernelStats{};
  if (p_.k > 64) {
    propagate_big();
    return;
  }
  const bool flip = p_.metric == kCosine && !p_.cosine_fixed;
  const int base = 2 * p_.trees;
  AKNN_HIP_TRY(hipMemsetAsync(ncand_, 0, sizeof(unsigned long long) * kCountWords * std::max(p_.props, 1), stream_));
  round_screened_.assign(std::max(p_.props, 0), 0);
  round_form_.assign(std::max(p_.props, 0), 0);
  // Screened or exact rounds (the graph is the same either way).  The first
  // calculate() on a data set runs round 0 screened and round 1 exact and
  // keeps the mode with the lower time per examined candidate for the other
  // rounds and later calls (clustered data can make the exact path's row
  // reuse win).  A round where more than a quarter of the candidates survived
  // the screen also switches the remaining rounds to the exact path.
  // with the principal-subspace screen every round screens with it (no probe)
  const bool pca = pca_.r > 0;
  const bool can_screen = propagation_screened() && !pca;
  rounds_from_graph_ = false;
  if (p_.graph_rounds && p_.props > 0 && (!can_screen || prop_mode_ >= 0) && propagate_graph(base)) return;
  bool screen = can_screen && prop_mode_ != 1;
  const bool probe = can_screen && prop_mode_ < 0 && p_.props >= 2;
  for (int r = 0; r < p_.props; ++r) {
    if (can_screen && r > 0) {
      unsigned long long c[2] = {0, 0};
      round_counts(r - 1, c);
      if (probe && r == 2) {
        float ms[2] = {0, 0};
        unsigned long long c0[2] = {0, 0};
        round_counts(0, c0);
        AKNN_HIP_TRY(hipEventElapsedTime(&ms[0], launch_ev(base), launch_ev(base + 1)));
        AKNN_HIP_TRY(hipEventElapsedTime(&ms[1], launch_ev(base + 2), launch_ev(base + 3)));
        const double per_s = ms[0] / std::max<double>(1.0, static_cast<double>(c0[0]));
        const double per_e = ms[1] / std::max<double>(1.0, static_cast<double>(c[0]));
        prop_mode_ = per_s <= per_e ? 0 : 1;
        screen = prop_mode_ == 0;
      } else if (probe && r == 1) {
        screen = false;  // the exact probe round
      }
      if (round_screened_[r - 1] && c[0] > 0 && 4 * c[1] > c[0]) screen = false;
    }
    round_screened_[r] = pca ? 2 : (screen ? 1 : 0);
    // buffers alternate each round: round r reads buffer cur_ (= c0 ^ (r & 1))
    const PropagateArgs a = round_args(p_, kmetric(), flip, screen, pca_, pca_proj_, pca_e_, x_, d_, ld_, xh_, ldh_, eps_, q8_ready_ ? q8_ : nullptr, ld8_, s8_,
                                       a8_, e8_, norms_, perm_, n_, gids_, gvals_, cur_, r, ncand_, changed_);
    AKNN_HIP_TRY(hipEventRecord(launch_ev(base + 2 * r), stream_));
    round_form_[r] = static_cast<uint8_t>(launch_propagate(a, stream_));
    AKNN_HIP_TRY(hipEventRecord(launch_ev(base + 2 * r + 1), stream_));
    cur_ ^= 1;  // buffer swap (.cu:1969-1975)
  }
  if (probe && prop_mode_ < 0) {
    // two rounds only: decide from the screened round 0 and the exact round 1
    unsigned long long c[4] = {0, 0, 0, 0};
    float ms[2] = {0, 0};
    round_counts(0, c);
    round_counts(1, c + 2);
    AKNN_HIP_TRY(hipEventElapsedTime(&ms[0], launch_ev(base), launch_ev(base + 1)));
    AKNN_HIP_TRY(hipEventElapsedTime(&ms[1], launch_ev(base + 2), launch_ev(base + 3)));
    const double per_s = ms[0] / std::max<double>(1.0, static_cast<double>(c[0]));
    const double per_e = ms[1] / std::max<double>(1.0, static_cast<double>(c[2]));
    prop_mode_ = per_s <= per_e ? 0 : 1;
  }
  stats_propagate.launches = p_.props;
}

// The P rounds replayed from one captured graph: the round mode is fixed (the
// probe of the first calculate() decided it, or there is no screen), so the
// launches and their buffers are known in advance.  The counters are cleared
// before the launch, outside the graph.  Returns false if capture is not
// possible (the caller then launches the rounds one by one).
bool Engine::propagate_graph(int base) {
  const bool pca = pca_.r > 0;
  const bool can_screen = propagation_screened() && !pca;
  const bool screen = can_screen && prop_mode_ == 0;
  const int mode = pca ? 2 + pca_.r : (screen ? 1 : 0);
  const bool flip = p_.metric == kCosine && !p_.cosine_fixed;
  if (pca && round_graph_sigma_ != pca_.sigma) drop_round_graph();  // sigma is a kernel argument
  if (!round_exec_ || round_graph_mode_ != mode) {
    drop_round_graph();
    hipGraph_t graph = nullptr;
    graph_form_.assign(p_.props, 0);
    AKNN_HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      for (int r = 0; r < p_.props; ++r)
        graph_form_[r] = static_cast<uint8_t>(launch_propagate(round_args(p_, kmetric(), flip, screen, pca_, pca_proj_, pca_e_, x_, d_, ld_, xh_, ldh_, eps_, q8_ready_ ? q8_ : nullptr, ld8_, s8_, a8_,
                                    e8_, norms_, perm_, n_, gids_, gvals_, (cur_ + r) & 1, r, ncand_, changed_),
                         stream_));
    } catch (...) {
      (void)hipStreamEndCapture(stream_, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    AKNN_HIP_TRY(hipStreamEndCapture(stream_, &graph));
    const hipError_t ie = hipGraphInstantiate(&round_exec_, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      (void)hipGetLastError();  // the rounds run one by one instead
      round_exec_ = nullptr;
      return false;
    }
    round_graph_mode_ = mode;
    round_graph_sigma_ = pca ? pca_.sigma : 0.0;
  }
  AKNN_HIP_TRY(hipEventRecord(launch_ev(base), stream_));
  AKNN_HIP_TRY(hipGraphLaunch(round_exec_, stream_));
  AKNN_HIP_TRY(hipEventRecord(launch_ev(base + 1), stream_));
  round_screened_.assign(p_.props, pca ? 2 : (screen ? 1 : 0));
  round_form_ = graph_form_;
  cur_ ^= p_.props & 1;
  rounds_from_graph_ = true;
  stats_propagate.launches = p_.props;
  return true;
}

void Engine::calculate() {
  check_shapes(p_, n_, d_);  // options may have changed the path since set_data
  AKNN_HIP_TRY(hipSetDevice(device_));
  cur_ = 0;
  draws_.reset(p_.seed);
  const bool sketch = p_.sketch || leaf_mfma_active();
  if (sketch && !xh_) {
    xh_ = alloc<_Float16>(static_cast<size_t>(n_) * ldh_);
    eps_ = alloc<float>(n_ + 1);  // [n]: the largest radius
    nh_ = alloc<float>(n_ + 1);   // [n]: the largest sketch norm
    // leaf re-scored, forest undecided, leaf screened: kSlotWords each
    leaf_rescored_ = alloc<unsigned long long>(3 * static_cast<size_t>(kSlotWords));
  }
  const bool prop8 = propagation_screened() && p_.sketch_int8 && p_.metric == kEuclidean && p_.props > 0;
  const bool forest8 = p_.sketch && p_.forest_int8 && p_.split == 0;  // the sign forest's side test
  const bool sketch8 = prop8 || forest8;
  if (sketch8 && !q8_) {
    q8_ = alloc<int8_t>(static_cast<size_t>(n_) * ld8_);
    s8_ = alloc<float>(n_);
    a8_ = alloc<int32_t>(n_);
    e8_ = alloc<float>(n_);
  }
  if (forest8 && !se8_) se8_ = alloc<float2>(n_);
  fb_.xh = p_.sketch ? xh_ : nullptr;
  fb_.ldh = ldh_;
  fb_.eps = p_.sketch ? eps_ : nullptr;
  fb_.nh = p_.sketch ? nh_ : nullptr;
  fb_.amb_count = p_.sketch ? leaf_rescored_ + kSlotWords : nullptr;
  fb_.q8 = forest8 ? q8_ : nullptr;
  fb_.ld8 = ld8_;
  fb_.se8 = forest8 ? se8_ : nullptr;
  AKNN_HIP_TRY(hipEventRecord(ev(0), stream_));
  // Router inference reads only the uploaded points: it runs on the side
  // stream from the start of the build (inside the timed bracket), beside
  // normalization, and the median forest waits for its bins.
  routers_launched_ = false;
  if (p_.split == 1 && !routers_.empty() && p_.router_trees > 0) {
    const int nr = static_cast<int>(routers_.size());
    for (Router* r : routers_) {
      if (r->device() != device_ || r->d() != d_) throw std::invalid_argument("router does not fit the data");
      if (r->bin_bits() != routers_[0]->bin_bits()) throw std::invalid_argument("routers must share bin_bits");
    }
    if (!bins_ || bins_cap_ < nr) {
      bins_ = alloc<int32_t>(static_cast<size_t>(n_) * nr);
      bins_cap_ = nr;
    }
    AKNN_HIP_TRY(hipStreamWaitEvent(aux_, ev(0), 0));  // after the previous build's readers of the bins
    for (int i = 0; i < nr; ++i) routers_[i]->assign(x_raw_, n_, ld_, bins_ + static_cast<int64_t>(i) * n_, aux_);
    AKNN_HIP_TRY(hipEventRecord(router_ev_, aux_));
    routers_launched_ = true;
  }
  if (p_.mean_fp64 && !mean64_) mean64_ = alloc<double>(static_cast<size_t>((n_ + 2047) / 2048) * d_);
  // the uploaded points stay as they are: normalization writes x_ from them
  // (and the FP16 sketch in the same row pass)
  SketchOut sko;
  sko.l1 = kmetric() == kTaxicab || kmetric() == kTaxicabL1;
  sko.xh = xh_;
  sko.ldh = ldh_;
  sko.eps = eps_;
  sko.nh = nh_;
  launch_normalize(x_raw_, x_, n_, d_, ld_, mean_, gmax_, p_.scale_all != 0, p_.mean_fp64 ? mean64_ : nullptr,
                   sketch ? &sko : nullptr, stream_);
  if (norms_) launch_norms(x_, n_, d_, ld_, norms_, stream_);
  if (sketch) {
    launch_max_nonneg(eps_, n_, eps_ + n_, stream_);
    launch_max_nonneg(nh_, n_, nh_ + n_, stream_);
  }
  if (sketch8) launch_sketch8(x_, n_, d_, ld_, q8_, ld8_, s8_, a8_, e8_, forest8 ? se8_ : nullptr, stream_);
  q8_ready_ = sketch8;  // the int8 rows are this build's
  pca_ = PcaBasis{};
  // This build's principal subspace: the sample moments on the device, then
  // the eigen-solve on a host thread while the device builds the forest and
  // the leaf lists (only the propagation rounds read the basis); every
  // point's projection and radius are made just before the rounds.
  std::thread pca_thread;
  PcaBasis pca_next;
  std::exception_ptr pca_error;
  if (p_.pca_screen && p_.metric == kEuclidean && p_.props > 0 && d_ >= 16 && d_ <= 256) {
    if (!pca_mom_) {
      pca_mom_ = alloc<float>(static_cast<size_t>(33) * d_ * d_);
      pca_p_ = alloc<float>(static_cast<size_t>(32) * d_);
      pca_proj_ = alloc<float>(static_cast<size_t>(n_) * 32);
      pca_e_ = alloc<float>(n_ + 1);  // [n]: the largest radius
    }
    if (launch_pca_moments(x_, n_, d_, ld_, pca_mom_, stream_)) {
      float* mh = static_cast<float*>(pinned(4, sizeof(float) * static_cast<size_t>(d_) * d_));
      AKNN_HIP_TRY(hipMemcpyAsync(mh, pca_mom_, sizeof(float) * static_cast<size_t>(d_) * d_, hipMemcpyDeviceToHost,
                                  stream_));
      AKNN_HIP_TRY(hipEventRecord(pca_ev_, stream_));
      pca_thread = std::thread([&, mh] {
        try {
          AKNN_HIP_TRY(hipSetDevice(device_));
          AKNN_HIP_TRY(hipEventSynchronize(pca_ev_));
          pca_next = pca_basis_from_moments(mh, d_);
        } catch (...) {
          pca_error = std::current_exception();
        }
      });
    }
  }
  auto pca_finish = [&] {
    if (!pca_thread.joinable()) return;
    pca_thread.join();
    if (pca_error) std::rethrow_exception(pca_error);
    pca_ = std::move(pca_next);
    if (pca_.r > 0) {
      AKNN_HIP_TRY(hipMemcpyAsync(pca_p_, pca_.p.data(), sizeof(float) * pca_.p.size(), hipMemcpyHostToDevice,
                                  stream_));
      if (p_.pca_fp16)
        launch_pca_project(x_, n_, d_, ld_, pca_p_, pca_.r, nullptr, pca_e_, stream_,
                           reinterpret_cast<_Float16*>(pca_proj_));
      else
        launch_pca_project(x_, n_, d_, ld_, pca_p_, pca_.r, pca_proj_, pca_e_, stream_);
      launch_max_nonneg(pca_e_, n_, pca_e_ + n_, stream_);
    }
  };
  AKNN_HIP_TRY(hipEventRecord(ev(1), stream_));
  try {
    if (p_.split == 1)
      build_forest_median();
    else
      build_forest();
    AKNN_HIP_TRY(hipEventRecord(ev(2), stream_));
    initial_knn();
    AKNN_HIP_TRY(hipEventRecord(ev(3), stream_));
  } catch (...) {
    if (pca_thread.joinable()) pca_thread.join();
    throw;
  }
  pca_finish();
  propagate();
  AKNN_HIP_TRY(hipEventRecord(ev(6), stream_));
  AKNN_HIP_TRY(hipEventSynchronize(ev(6)));
  float a = 0, b = 0, c = 0, dd = 0, e = 0;
  AKNN_HIP_TRY(hipEventElapsedTime(&a, ev(0), ev(1)));
  AKNN_HIP_TRY(hipEventElapsedTime(&b, ev(1), ev(2)));
  AKNN_HIP_TRY(hipEventElapsedTime(&c, ev(2), ev(3)));
  AKNN_HIP_TRY(hipEventElapsedTime(&dd, ev(3), ev(6)));
  AKNN_HIP_TRY(hipEventElapsedTime(&e, ev(0), ev(6)));
  times_ = StageTimes{a, b, c, dd, e};
  const bool mfma = leaf_mfma_active();
  for (int64_t l = 0; l < stats_leaf_knn.launches; ++l) {
    float ms = 0;
    AKNN_HIP_TRY(hipEventElapsedTime(&ms, launch_ev(2 * l), launch_ev(2 * l + 1)));
    stats_leaf_knn.total_ms += ms;
    if (!leaf_all_used_ && mfma && tree_screened_[l]) {
      stats_leaf_knn.screened_ms += ms;
      stats_leaf_knn.screened_launches += 1;
    }
  }
  if (leaf_all_used_) leaf_all_stats(mfma);
  std::vector<unsigned long long> nc(2 * static_cast<size_t>(std::max(p_.props, 1)));
  for (int r = 0; r < p_.props; ++r) round_counts(r, &nc[2 * r]);
  stats_propagate.examined = stats_propagate.rescored = 0;
  if (rounds_from_graph_ && p_.props > 0) {
    float ms = 0;
    AKNN_HIP_TRY(hipEventElapsedTime(&ms, launch_ev(2 * p_.trees), launch_ev(2 * p_.trees + 1)));
    stats_propagate.total_ms = ms;
  }
  for (int r = 0; r < p_.props; ++r) {
    float ms = 0;
    if (!rounds_from_graph_) {
      AKNN_HIP_TRY(hipEventElapsedTime(&ms, launch_ev(2 * p_.trees + 2 * r), launch_ev(2 * p_.trees + 2 * r + 1)));
    }
    stats_propagate.total_ms += ms;
    const double k = p_.k;
    // per point: own row (8k), k neighbour id rows (4k^2), x_i (4D), output
    // row (8k); per unique new candidate its row (4D), or with a screen its
    // screen row plus 4D for each exactly re-scored one: the FP16 sketch row
    // (2 ldh; COSINE also its norm), the int8 row with its three scalars
    // (scale, radius and sum of squares or, COSINE, the norm), or the FP32
    // (4r) or FP16 (2r) projection.  The FP16 sketch and projection screens
    // bound every candidate's radius by the largest, read once per point.
    double per_point = 16.0 * k + 4.0 * k * k + 4.0 * d_;
    double cand = 4.0 * d_ * nc[2 * r];
    if (round_screened_[r] && round_form_[r] > 0) {
      const int f = round_form_[r];
      const double row = f == 1   ? 2.0 * ldh_ + (p_.metric == kCosine ? 4.0 : 0.0)
                         : f == 2 ? ld8_ + 12.0
                         : f == 3 ? 4.0 * pca_.r
                                  : 2.0 * pca_.r;
      per_point += row + 8.0;  // own screen row, own radius and the radius bound
      cand = row * nc[2 * r] + 4.0 * d_ * nc[2 * r + 1];
    }
    stats_propagate.units += static_cast<double>(n_) * per_point + cand;
    stats_propagate.examined += static_cast<double>(nc[2 * r]);
    stats_propagate.rescored += static_cast<double>(round_screened_[r] && round_form_[r] > 0 ? nc[2 * r + 1] : nc[2 * r]);
  }
  stats_forest.launches = 0;
  for (int t = 0; t < p_.trees; ++t) stats_forest.launches += iters_[t];
  stats_forest.total_ms = b;
  // leaf k-NN flops: 3D per pair (subtract + multiply-add), pairs = every
  // owner against its start node minus itself
  if (!(p_.metric == kCosine && !p_.cosine_fixed)) {
    std::vector<Leaf> host;
    double pairs = 0, screened = 0;
    for (int t = 0; t < p_.trees; ++t) {
      host.resize(nleaves_[t]);
      AKNN_HIP_TRY(hipMemcpy(host.data(), leaves_ + t * n_, sizeof(Leaf) * nleaves_[t], hipMemcpyDeviceToHost));
      for (const Leaf& l : host) {
        pairs += static_cast<double>(l.size) * (l.rcount - 1);
        if (mfma && tree_screened_[t]) screened += static_cast<double>(l.size) * (l.rcount - 1);
      }
    }
    stats_leaf_knn.units = pairs * 3.0 * d_;
    stats_leaf_knn.screened = screened;
    if (mfma) {
      unsigned long long r = 0;
      slot_sums(leaf_rescored_, 1, &r);
      stats_leaf_knn.rescored = static_cast<double>(r);
    }
  }
}

void Engine::get_graph(int32_t* ids, float* vals) const {
  const size_t nk = static_cast<size_t>(n_) * p_.k;
  AKNN_HIP_TRY(hipMemcpy(ids, gids_[cur_], nk * sizeof(int32_t), hipMemcpyDeviceToHost));
  AKNN_HIP_TRY(hipMemcpy(vals, gvals_[cur_], nk * sizeof(float), hipMemcpyDeviceToHost));
}

void Engine::get_normalized(float* x) const {
  AKNN_HIP_TRY(hipMemcpy2D(x, d_ * sizeof(float), x_, ld_ * sizeof(float), d_ * sizeof(float), n_,
                           hipMemcpyDeviceToHost));
}

namespace {
// Device scratch for one call, freed on every exit.
struct DeviceScratch {
  std::vector<void*> p;
  template <typename T>
  T* get(size_t count) {
    void* q = nullptr;
    AKNN_HIP_TRY(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
    p.push_back(q);
    return static_cast<T*>(q);
  }
  ~DeviceScratch() {
    for (void* q : p) (void)hipFree(q);
  }
};
}  // namespace

void Engine::exact_knn(const int64_t* queries, int64_t nq, int kq, int32_t* ids, double* dist) {
  if (nq < 0 || kq < 1 || kq > 64 || kq >= n_) throw std::invalid_argument("exact k-NN needs 1 <= k <= 64, k < N");
  for (int64_t q = 0; q < nq; ++q)
    if (queries[q] < 0 || queries[q] >= n_) throw std::invalid_argument("query index out of range");
  if (nq == 0) return;
  AKNN_HIP_TRY(hipSetDevice(device_));
  DeviceScratch sc;
  int64_t* dq = sc.get<int64_t>(nq);
  int32_t* di = sc.get<int32_t>(static_cast<size_t>(nq) * kq);
  double* dd = sc.get<double>(static_cast<size_t>(nq) * kq);
  double* sq = nullptr;
  if (p_.metric == kCosine) {
    sq = sc.get<double>(n_);
    launch_sq64(x_, n_, d_, ld_, sq, stream_);
  }
  AKNN_HIP_TRY(hipMemcpyAsync(dq, queries, sizeof(int64_t) * nq, hipMemcpyHostToDevice, stream_));
  launch_exact_knn(p_.metric, x_, n_, d_, ld_, sq, dq, nq, kq, di, dd, stream_);
  AKNN_HIP_TRY(hipMemcpyAsync(ids, di, sizeof(int32_t) * nq * kq, hipMemcpyDeviceToHost, stream_));
  AKNN_HIP_TRY(hipMemcpyAsync(dist, dd, sizeof(double) * nq * kq, hipMemcpyDeviceToHost, stream_));
  AKNN_HIP_TRY(hipStreamSynchronize(stream_));
}

double Engine::recall(const int64_t* queries, int64_t nq) {
  const int k = p_.k;
  for (int64_t q = 0; q < nq; ++q)
    if (queries[q] < 0 || queries[q] >= n_) throw std::invalid_argument("query index out of range");
  if (nq <= 0) return 0.0;
  AKNN_HIP_TRY(hipSetDevice(device_));
  DeviceScratch sc;
  int64_t* dq = sc.get<int64_t>(nq);
  int32_t* di = sc.get<int32_t>(static_cast<size_t>(nq) * k);
  double* dd = sc.get<double>(static_cast<size_t>(nq) * k);
  double* kth = sc.get<double>(nq);
  unsigned long long* hits = sc.get<unsigned long long>(1);
  double* sq = nullptr;
  if (p_.metric == kCosine) {
    sq = sc.get<double>(n_);
    launch_sq64(x_, n_, d_, ld_, sq, stream_);
  }
  AKNN_HIP_TRY(hipMemcpyAsync(dq, queries, sizeof(int64_t) * nq, hipMemcpyHostToDevice, stream_));
  AKNN_HIP_TRY(hipMemsetAsync(hits, 0, sizeof(unsigned long long), stream_));
  launch_exact_knn(p_.metric, x_, n_, d_, ld_, sq, dq, nq, k, di, dd, stream_);
  // the k-th exact distance of query q is dd[q * k + k - 1]
  AKNN_HIP_TRY(hipMemcpy2DAsync(kth, sizeof(double), dd + (k - 1), sizeof(double) * k, sizeof(double), nq,
                                hipMemcpyDeviceToDevice, stream_));
  launch_recall(p_.metric, x_, d_, ld_, sq, dq, nq, kth, gids_[cur_], k, hits, stream_);
  unsigned long long h = 0;
  AKNN_HIP_TRY(hipMemcpyAsync(&h, hits, sizeof(h), hipMemcpyDeviceToHost, stream_));
  AKNN_HIP_TRY(hipStreamSynchronize(stream_));
  return static_cast<double>(h) / (static_cast<double>(nq) * k);
}

void Engine::get_forest(int tree, int32_t* perm, int32_t* rs, int32_t* rc) {
  if (tree < 0 || tree >= p_.trees || tree >= static_cast<int>(nleaves_.size()))
    throw std::invalid_argument("no such tree (run calculate first)");
  if (!rs_) {
    rs_ = alloc<int32_t>(static_cast<size_t>(n_) * p_.trees);
    rc_ = alloc<int32_t>(static_cast<size_t>(n_) * p_.trees);
  }
  launch_leaf_ranges(perm_ + tree * n_, leaves_ + tree * n_, nleaves_[tree], rs_ + tree * n_, rc_ + tree * n_,
                     stream_);
  AKNN_HIP_TRY(hipStreamSynchronize(stream_));
  AKNN_HIP_TRY(hipMemcpy(perm, perm_ + tree * n_, n_ * sizeof(int32_t), hipMemcpyDeviceToHost));
  AKNN_HIP_TRY(hipMemcpy(rs, rs_ + tree * n_, n_ * sizeof(int32_t), hipMemcpyDeviceToHost));
  AKNN_HIP_TRY(hipMemcpy(rc, rc_ + tree * n_, n_ * sizeof(int32_t), hipMemcpyDeviceToHost));
}

}  // namespace aknn
