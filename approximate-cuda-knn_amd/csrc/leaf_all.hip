// The initial leaf k-NN of every tree in one persistent launch.
//
// Reference: findInitialStateOfAproximated{Euclidean,Taxicab,Cosine}KNN
// (euclideandistancematrixgpu.cu:107-415, :417-712, :714-1031), where one
// kernel takes each point's candidates from all trees (.cu:160-191).  Here
// every tree merges its start-node candidates into the graph rows the earlier
// trees left (the selection decomposes over trees: the first k of stubs and
// all candidates by (key, id)), so the trees run in order inside one
// cooperative launch, separated by a grid barrier; a tree's work items (big
// leaves cut into 64-owner items, runs of small ones grouped) of uneven sizes
// are taken from eight queues, one per XCD (consecutive items share candidate
// rows, so each XCD takes a contiguous run into its own L2), and a wave whose
// queue is empty takes from the others.  The four waves of a workgroup take
// items independently; workgroups meet at the barriers.
//
// Per tree one of two item bodies runs: the exact FP32 one (leaf_exact.hpp)
// or the MFMA-screened one (leaf_screen.hpp), which store the same graph.
// Tree 0 meets stub rows (nothing to screen against) and runs exact; the
// next trees run screened while the screen pays.  After a screened tree the
// last workgroup to reach the barrier (every other one has finished that
// tree) reads the tree's re-scored and screened pair counts and switches the
// remaining trees to the exact body when more than 30% of the pairs were
// re-scored.  Each barrier wait is bounded (a wall-clock limit sets an error
// flag and lets the wave go on; the host then reports the build as failed).
// When the runtime refuses the cooperative launch, nothing has run and the
// caller falls back to one launch per tree.
#include <vector>

#include "leaf_exact.hpp"
#include "leaf_screen.hpp"

namespace aknn {

namespace {

// One tree's arguments of the two item bodies (work items, leaf order and
// start ranges of that tree), read from device memory by the workgroups.
struct TreeLeafArgs {
  leafs::MfmaLeafArgs m;
  leafx::ExactArgs e;
};

struct AllLeafArgs {
  const TreeLeafArgs* targs;  // [trees]
  int trees;
  int screen;  // trees after the first may run screened
  unsigned long long* rescored;  // the screened body's spread counters
  unsigned long long* screened;
  LeafAllState st;
  unsigned long long spin_limit;  // wall-clock ticks a barrier wait may take
};

// Sum of a spread counter's slots.
__device__ unsigned long long slot_total(const unsigned long long* c) {
  unsigned long long s = 0;
  for (int i = 0; i < kCountSlots; ++i) s += __hip_atomic_load(c + i * kCountStride, __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_AGENT);
  return s;
}

// Grid barrier after tree t.  Each workgroup releases its rows with its
// arrival (one agent-scope release) and acquires the others' once the last
// one has arrived (one agent-scope acquire); the wait itself polls with
// relaxed loads, which keep the caches of the workgroups still working.  The
// workgroup that arrives last decides tree t + 1's body and stamps the clock,
// then releases the others.
__device__ void tree_barrier(const AllLeafArgs& a, int t) {
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned* bar = a.st.bar;
    const unsigned gen = __hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned arrived = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) + 1u;
    if (arrived == gridDim.x) {
      __atomic_thread_fence(__ATOMIC_ACQUIRE);  // every other workgroup's rows and counts
      if (t + 1 < a.trees) {
        int next = 0;
        if (a.screen) {
          const int cur = __hip_atomic_load(a.st.mode + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (t == 0) {
            next = 1;
          } else if (cur == 1) {
            // this tree's increments of the re-scored and screened counters
            const unsigned long long res = slot_total(a.rescored), scr = slot_total(a.screened);
            const unsigned long long p0 = __hip_atomic_load(a.st.prev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long p1 = __hip_atomic_load(a.st.prev + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long dres = res - p0, dscr = scr - p1;
            __hip_atomic_store(a.st.prev, res, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(a.st.prev + 1, scr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            next = dscr > 0 && 10 * dres > 3 * dscr ? 0 : 1;
          }
        }
        __hip_atomic_store(a.st.mode + t + 1, next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      __hip_atomic_store(a.st.stamps + t + 1, static_cast<unsigned long long>(wall_clock64()), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(bar + 1, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      const unsigned long long t0 = static_cast<unsigned long long>(wall_clock64());
      while (__hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        __builtin_amdgcn_s_sleep(8);
        if (static_cast<unsigned long long>(wall_clock64()) - t0 > a.spin_limit) {
          atomicExch(a.st.error, 1);
          break;
        }
      }
      __atomic_thread_fence(__ATOMIC_ACQUIRE);  // the rows of the workgroups that arrived before
    }
  }
  __syncthreads();
}

// Next work item of tree t for a workgroup on XCD xq: its own queue's run
// first, then the other queues' in turn (*j tracks the queue); -1 when all
// of the tree's items are taken.
__device__ __forceinline__ int next_item(const AllLeafArgs& a, int t, int ni, int xq, int* j) {
  for (; *j < 8; ++*j) {
    const int q = (xq + *j) & 7;
    const int lo = static_cast<int>(static_cast<int64_t>(ni) * q / 8);
    const int hi = static_cast<int>(static_cast<int64_t>(ni) * (q + 1) / 8);
    int it = hi;
    if ((threadIdx.x & 63) == 0) {
      const unsigned got = atomicAdd(a.st.queue + t * 8 + q, 1u);
      it = got < static_cast<unsigned>(hi - lo) ? lo + static_cast<int>(got) : hi;
    }
    it = __shfl(it, 0);
    if (it < hi) return it;
  }
  return -1;
}

// Four waves per workgroup, each taking work items on its own with its own
// LDS; the workgroups, not the waves, meet at the tree barriers (a quarter
// of the cache maintenance).  2 waves per SIMD (1 for lists of 64): the LDS
// of one wave (the larger body's) allows no more.
constexpr int kGroupWaves = 4;

template <int METRIC, int KMAX, bool BIGD>
__global__ __launch_bounds__(64 * kGroupWaves) __attribute__((amdgpu_waves_per_eu(KMAX <= 32 ? 2 : 1, 2)))
void leaf_all_kernel(AllLeafArgs a) {
  constexpr bool kScreenable = METRIC == kEuclidean || METRIC == kCosine;
  __shared__ union Lds {
    leafx::ExactLds e;
    leafs::ScreenLds m;
  } sm[kGroupWaves];
  Lds& my = sm[threadIdx.x >> 6];
  if (blockIdx.x == 0 && threadIdx.x == 0)
    __hip_atomic_store(a.st.stamps, static_cast<unsigned long long>(wall_clock64()), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  const int xq = blockIdx.x & 7;  // workgroups are dealt to the XCDs in turn
  for (int t = 0; t < a.trees; ++t) {
    const int mode =
        t == 0 ? 0 : __hip_atomic_load(a.st.mode + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int j = 0;
    if (kScreenable && mode == 1) {
      const leafs::MfmaLeafArgs m = a.targs[t].m;
      for (int it = next_item(a, t, m.nleaves, xq, &j); it >= 0; it = next_item(a, t, m.nleaves, xq, &j))
        leafs::screen_item<METRIC, KMAX, false>(m, it, my.m);
    } else {
      const leafx::ExactArgs e = a.targs[t].e;
      for (int it = next_item(a, t, e.nleaves, xq, &j); it >= 0; it = next_item(a, t, e.nleaves, xq, &j))
        leafx::exact_item<METRIC, KMAX, BIGD, false>(e, it, my.e);
    }
    tree_barrier(a, t);
  }
}

template <int METRIC, int KMAX, bool BIGD>
bool launch_all(const AllLeafArgs& a, int device, hipStream_t s) {
  const void* f = reinterpret_cast<const void*>(&leaf_all_kernel<METRIC, KMAX, BIGD>);
  int per_cu = 0, cus = 0, coop = 0;
  AKNN_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 64 * kGroupWaves, 0));
  AKNN_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  AKNN_HIP_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device));
  const int grid = per_cu * cus / 8 * 8;
  if (!coop || grid < 8) return false;
  AllLeafArgs arg = a;
  void* params[] = {&arg};
  // a refused cooperative launch (too large a grid, or the runtime cannot
  // give this process one now, e.g. while another process holds the device)
  // launches nothing: the caller runs the trees one launch each
  const hipError_t e = hipLaunchCooperativeKernel(f, dim3(grid), dim3(64 * kGroupWaves), params, 0, s);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return true;
}

template <int METRIC, bool BIGD>
bool launch_k(const AllLeafArgs& a, int k, int device, hipStream_t s) {
  if (k <= 8) return launch_all<METRIC, 8, BIGD>(a, device, s);
  if (k <= 16) return launch_all<METRIC, 16, BIGD>(a, device, s);
  if (k <= 20) return launch_all<METRIC, 20, BIGD>(a, device, s);
  if (k <= 24) return launch_all<METRIC, 24, BIGD>(a, device, s);
  if (k <= 32) return launch_all<METRIC, 32, BIGD>(a, device, s);
  if (k <= 64) return launch_all<METRIC, 64, BIGD>(a, device, s);
  throw HipError("leaf k-NN supports k <= 64");
}

}  // namespace

size_t leaf_all_targs_bytes(int trees) { return sizeof(TreeLeafArgs) * static_cast<size_t>(trees); }

bool launch_leaf_all(int metric, bool screen, const float* x, int d, int ld, const _Float16* xh, int ldh,
                     const float* eps, const float* eps_max, const float* nh, const float* norms, const int32_t* perm,
                     const Leaf* items, const int64_t* item_off, int trees, int64_t n, int k, int32_t* ids,
                     float* vals, const int32_t* rs, const int32_t* rc, unsigned long long* rescored,
                     unsigned long long* screened, const LeafAllState& st, int device, hipStream_t s) {
  if (trees < 1) return true;
  screen = screen && (metric == kEuclidean || metric == kCosine) && trees > 1;
  if (screen && (xh == nullptr || eps == nullptr || eps_max == nullptr || nh == nullptr || rescored == nullptr ||
                 screened == nullptr || ldh % 8 != 0 || ldh < d))
    throw HipError("leaf k-NN: the screened path needs the sketch and its counters");
  if (metric == kCosine && norms == nullptr) throw HipError("leaf k-NN: COSINE needs the norms");
  // each tree's arguments of both bodies (the per-tree kernels' arguments)
  std::vector<TreeLeafArgs> ta(trees);
  leafs::MfmaLeafArgs m0{};
  if (screen) m0 = leafs::screen_args(metric, x, d, ld, xh, ldh, eps, eps_max, nh, norms, k, ids, vals, rescored,
                                      screened);
  for (int t = 0; t < trees; ++t) {
    const int ni = static_cast<int>(item_off[t + 1] - item_off[t]);
    const int64_t tn = static_cast<int64_t>(t) * n;
    ta[t].e = leafx::ExactArgs{x, d, ld, norms, perm + tn, items + item_off[t], ni, 1, rs + tn, rc + tn, k, ids, vals};
    ta[t].m = m0;
    ta[t].m.perm = perm + tn;
    ta[t].m.leaves = items + item_off[t];
    ta[t].m.nleaves = ni;
    ta[t].m.chunks_per_leaf = 1;
    ta[t].m.rsp = rs + tn;
    ta[t].m.rcp = rc + tn;
  }
  AKNN_HIP_TRY(hipMemcpyAsync(st.targs, ta.data(), leaf_all_targs_bytes(trees), hipMemcpyHostToDevice, s));
  AllLeafArgs a{};
  a.targs = static_cast<const TreeLeafArgs*>(st.targs);
  a.trees = trees;
  a.screen = screen ? 1 : 0;
  a.rescored = rescored;
  a.screened = screened;
  a.st = st;
  int khz = 0;
  AKNN_HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
  a.spin_limit = static_cast<unsigned long long>(std::max(khz, 1)) * 1000ull * 60ull;  // one minute
  AKNN_HIP_TRY(hipMemsetAsync(st.queue, 0, sizeof(unsigned) * 8 * static_cast<size_t>(trees), s));
  AKNN_HIP_TRY(hipMemsetAsync(st.bar, 0, sizeof(unsigned) * 2, s));
  AKNN_HIP_TRY(hipMemsetAsync(st.mode, 0, sizeof(int32_t) * static_cast<size_t>(trees), s));
  AKNN_HIP_TRY(hipMemsetAsync(st.prev, 0, sizeof(unsigned long long) * 2, s));
  AKNN_HIP_TRY(hipMemsetAsync(st.error, 0, sizeof(int), s));
  const bool big = d > 256;
  switch (metric) {
    case kEuclidean:
      return big ? launch_k<kEuclidean, true>(a, k, device, s) : launch_k<kEuclidean, false>(a, k, device, s);
    case kTaxicab:
      return big ? launch_k<kTaxicab, true>(a, k, device, s) : launch_k<kTaxicab, false>(a, k, device, s);
    case kTaxicabL1:
      return big ? launch_k<kTaxicabL1, true>(a, k, device, s) : launch_k<kTaxicabL1, false>(a, k, device, s);
    case kCosine:
      return big ? launch_k<kCosine, true>(a, k, device, s) : launch_k<kCosine, false>(a, k, device, s);
    default:
      throw HipError("unknown distance");
  }
}

}  // namespace aknn
