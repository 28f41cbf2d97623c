// Shared definitions of the MI355X k-NN graph builder (host and device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace aknn {

// Distance codes, as euclideandistancematrixgpu.h:16-18.
constexpr int kEuclidean = 0;
constexpr int kTaxicab = 1;
constexpr int kCosine = 2;
// Internal code of TAXICAB without the square root (option taxicab_l1): the
// reference stores sqrtf of the L1 total (euclideandistancematrixgpu.cu:568,
// :1477); this fixed mode keys and stores the L1 total itself.
constexpr int kTaxicabL1 = 3;

// Stub neighbour written before the initial pass (euclideandistancematrixgpu.cu:16-17).
constexpr int kStubId = -1;
constexpr float kStubDist = 100000.0f;

// One finished bucket of a tree: positions [start, start+size) of the leaf
// array, and the range [rstart, rstart+rcount) of the deepest ancestor-or-self
// node with at least k+1 points (the "start node" of every member).
struct Leaf {
  int32_t start;
  int32_t size;
  int32_t rstart;
  int32_t rcount;
};

// One bucket still being split.  astart indexes the compacted array of active
// elements; fstart is the bucket's first position in the final leaf array.
struct Bucket {
  int32_t astart;
  int32_t size;
  int32_t fstart;
  int32_t q1;
  int32_t q2;
  int32_t rstart;
  int32_t rcount;
  int32_t rbefore;   // RIGHT elements before astart (filled per level)
  int32_t nright;    // RIGHT elements of the bucket (filled per level)
  int32_t tree;      // tree of the bucket within a batch of trees built together (0 for one tree)
  int32_t pad[6];
};
static_assert(sizeof(Bucket) == 64, "Bucket is one 64-byte record");

// Device-side control block of the per-tree level loop.
struct ForestCtl {
  int32_t nact[2];     // active elements at even/odd levels
  int32_t nbkt[2];     // active buckets at even/odd levels
  int32_t done;        // set by the first level that has no bucket to split
  int32_t iterations;  // reference loop iterations (levels with a split + 1)
  int32_t nleaves;     // leaves written so far
  int32_t total_right; // RIGHT elements of the current level
  int64_t active_sum;  // sum over levels of active elements (for byte accounting)
};

}  // namespace aknn

#define AKNN_HIP_TRY(expr)                                                            \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      throw aknn::HipError(std::string(#expr) + ": " + hipGetErrorString(e_) + " (" + \
                           __FILE__ + ":" + std::to_string(__LINE__) + ")");          \
    }                                                                                 \
  } while (0)

namespace aknn {
struct HipError {
  std::string what;
  explicit HipError(std::string w) : what(std::move(w)) {}
};
}  // namespace aknn
