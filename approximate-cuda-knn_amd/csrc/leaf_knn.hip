// Initial leaf k-NN for one tree (EUCLIDEAN, TAXICAB, COSINE with carried
// distances), merged into the graph rows in place.
//
// Reference: findInitialStateOfAproximated{Euclidean,Taxicab,Cosine}KNN
// (euclideandistancematrixgpu.cu:107-415, :417-712, :714-1031).  There, every
// point's candidates are the union over trees of its start node's points
// (.cu:160-191), scanned in chunks against the k entries carried so far and
// reduced by a histogram select.  When the stored value equals the selection
// key (EUCLIDEAN, TAXICAB, COSINE with distances carried) that loop keeps the
// first k of (stubs + all candidates) ordered by (key, id), so it decomposes
// exactly over trees: each tree's pass merges its start-node candidates into
// the running top-k (duplicates across trees carry identical keys and are
// dropped by id).
//
// MI355X mapping: one wave per (leaf, 64-owner chunk).  A leaf's points share
// one start node, so the wave streams that node's points in tiles of 64 and
// computes the 64x64 distance block as a register-blocked outer product (8x8
// pairs per lane, dimension chunks staged transposed in LDS).  Each pair's
// chain still runs over the dimensions in order, so every key is bit-exact.
// The block goes through LDS to the owner lanes, which keep their top-k in
// registers (seeded from the graph row, so later trees reject most pairs at
// the threshold test).
#include "leaf_exact.hpp"

namespace aknn {

namespace {

using leafx::kTile;

template <int METRIC, int KMAX, bool BIGD>
__global__ __launch_bounds__(64) void leaf_knn_kernel(leafx::ExactArgs a) {
  __shared__ leafx::ExactLds sm;
  leafx::exact_item<METRIC, KMAX, BIGD>(a, xcd_block(blockIdx.x, gridDim.x), sm);
}

template <int METRIC, int KMAX, bool BIGD>
void launch_t(const float* x, int d, int ld, const float* norms, const int32_t* perm, const Leaf* leaves, int nleaves,
              int max_leaf, int k, int32_t* ids, float* vals, const int32_t* rsp, const int32_t* rcp, hipStream_t s) {
  const int chunks = (max_leaf + kTile - 1) / kTile;
  const int64_t grid = xcd_grid(static_cast<int64_t>(nleaves) * chunks);
  const leafx::ExactArgs a{x, d, ld, norms, perm, leaves, nleaves, chunks, rsp, rcp, k, ids, vals};
  hipLaunchKernelGGL((leaf_knn_kernel<METRIC, KMAX, BIGD>), dim3(static_cast<unsigned>(grid)), dim3(kTile), 0, s, a);
  AKNN_HIP_TRY(hipGetLastError());
}

template <int METRIC, bool BIGD>
void launch_k(int k, const float* x, int d, int ld, const float* norms, const int32_t* perm, const Leaf* leaves,
              int nleaves, int max_leaf, int32_t* ids, float* vals, const int32_t* rsp, const int32_t* rcp,
              hipStream_t s) {
  if (k <= 8)
    launch_t<METRIC, 8, BIGD>(x, d, ld, norms, perm, leaves, nleaves, max_leaf, k, ids, vals, rsp, rcp, s);
  else if (k <= 16)
    launch_t<METRIC, 16, BIGD>(x, d, ld, norms, perm, leaves, nleaves, max_leaf, k, ids, vals, rsp, rcp, s);
  else if (k <= 20)
    launch_t<METRIC, 20, BIGD>(x, d, ld, norms, perm, leaves, nleaves, max_leaf, k, ids, vals, rsp, rcp, s);
  else if (k <= 24)
    launch_t<METRIC, 24, BIGD>(x, d, ld, norms, perm, leaves, nleaves, max_leaf, k, ids, vals, rsp, rcp, s);
  else if (k <= 32)
    launch_t<METRIC, 32, BIGD>(x, d, ld, norms, perm, leaves, nleaves, max_leaf, k, ids, vals, rsp, rcp, s);
  else if (k <= 64)
    launch_t<METRIC, 64, BIGD>(x, d, ld, norms, perm, leaves, nleaves, max_leaf, k, ids, vals, rsp, rcp, s);
  else
    throw HipError("leaf k-NN supports k <= 64");
}

// stubInitializationKernel (.cu:2110-2117)
__global__ __launch_bounds__(256) void stub_init_kernel(int32_t* __restrict__ ids, float* __restrict__ vals,
                                                        int64_t count) {
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < count; e += gridDim.x * 256ll) {
    ids[e] = kStubId;
    vals[e] = kStubDist;
  }
}

}  // namespace

void launch_stub_init(int32_t* ids, float* vals, int64_t count, hipStream_t s) {
  int64_t g = (count + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(stub_init_kernel, dim3(static_cast<unsigned>(g)), dim3(256), 0, s, ids, vals, count);
  AKNN_HIP_TRY(hipGetLastError());
}

void launch_leaf_knn(int metric, const float* x, int d, int ld, const float* norms, const int32_t* perm,
                     const Leaf* leaves, int nleaves, int max_leaf, int k, int32_t* ids, float* vals,
                     const int32_t* rsp, const int32_t* rcp, hipStream_t s) {
  if (nleaves <= 0) return;
  const bool big = d > 256;
  switch (metric) {
    case kEuclidean:
      big ? launch_k<kEuclidean, true>(k, x, d, ld, norms, perm, leaves, nleaves, max_leaf, ids, vals, rsp, rcp, s)
          : launch_k<kEuclidean, false>(k, x, d, ld, norms, perm, leaves, nleaves, max_leaf, ids, vals, rsp, rcp, s);
      break;
    case kTaxicab:
      big ? launch_k<kTaxicab, true>(k, x, d, ld, norms, perm, leaves, nleaves, max_leaf, ids, vals, rsp, rcp, s)
          : launch_k<kTaxicab, false>(k, x, d, ld, norms, perm, leaves, nleaves, max_leaf, ids, vals, rsp, rcp, s);
      break;
    case kTaxicabL1:
      big ? launch_k<kTaxicabL1, true>(k, x, d, ld, norms, perm, leaves, nleaves, max_leaf, ids, vals, rsp, rcp, s)
          : launch_k<kTaxicabL1, false>(k, x, d, ld, norms, perm, leaves, nleaves, max_leaf, ids, vals, rsp, rcp, s);
      break;
    case kCosine:
      big ? launch_k<kCosine, true>(k, x, d, ld, norms, perm, leaves, nleaves, max_leaf, ids, vals, rsp, rcp, s)
          : launch_k<kCosine, false>(k, x, d, ld, norms, perm, leaves, nleaves, max_leaf, ids, vals, rsp, rcp, s);
      break;
    default:
      throw HipError("unknown distance");
  }
}

}  // namespace aknn
