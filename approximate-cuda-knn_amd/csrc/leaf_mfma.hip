// Initial leaf k-NN for one tree with half-precision MFMA screening
// (EUCLIDEAN, COSINE with carried distances): the per-tree kernel over the
// item body of leaf_screen.hpp, and its launcher.
#include "leaf_screen.hpp"

namespace aknn {

namespace {

using namespace leafs;

template <int METRIC, int KMAX>
__global__ __launch_bounds__(64) void leaf_mfma_kernel(MfmaLeafArgs a) {
  __shared__ ScreenLds sm;
  screen_item<METRIC, KMAX>(a, xcd_block(blockIdx.x, gridDim.x), sm);
}

template <int METRIC, int KMAX>
void launch_k(const MfmaLeafArgs& a, hipStream_t s) {
  const int64_t grid = xcd_grid(static_cast<int64_t>(a.nleaves) * a.chunks_per_leaf);
  hipLaunchKernelGGL((leaf_mfma_kernel<METRIC, KMAX>), dim3(static_cast<unsigned>(grid)), dim3(kTile), 0, s, a);
  AKNN_HIP_TRY(hipGetLastError());
}

template <int METRIC>
void launch_m(const MfmaLeafArgs& a, int k, hipStream_t s) {
  if (k <= 8)
    launch_k<METRIC, 8>(a, s);
  else if (k <= 16)
    launch_k<METRIC, 16>(a, s);
  else if (k <= 20)
    launch_k<METRIC, 20>(a, s);
  else if (k <= 24)
    launch_k<METRIC, 24>(a, s);
  else if (k <= 32)
    launch_k<METRIC, 32>(a, s);
  else if (k <= 64)
    launch_k<METRIC, 64>(a, s);
  else
    throw HipError("leaf k-NN supports k <= 64");
}

}  // namespace

void launch_leaf_mfma(int metric, const float* x, int d, int ld, const _Float16* xh, int ldh, const float* eps,
                      const float* eps_max, const float* nh, const float* norms, const int32_t* perm, const Leaf* leaves, int nleaves,
                      int max_leaf, int k, int32_t* ids, float* vals, const int32_t* rsp, const int32_t* rcp,
                      unsigned long long* rescored, unsigned long long* screened, hipStream_t s) {
  if (nleaves <= 0) return;
  if (metric != kEuclidean && metric != kCosine) throw HipError("MFMA leaf screening supports EUCLIDEAN and COSINE");
  if (metric == kCosine && norms == nullptr) throw HipError("MFMA leaf screening: COSINE needs the norms");
  if (ldh % 8 != 0 || ldh < d) throw HipError("MFMA leaf screening: malformed sketch");
  MfmaLeafArgs a = screen_args(metric, x, d, ld, xh, ldh, eps, eps_max, nh, norms, k, ids, vals, rescored, screened);
  a.perm = perm;
  a.leaves = leaves;
  a.rsp = rsp;
  a.rcp = rcp;
  a.nleaves = nleaves;
  a.chunks_per_leaf = (max_leaf + kTile - 1) / kTile;
  if (metric == kEuclidean)
    launch_m<kEuclidean>(a, k, s);
  else
    launch_m<kCosine>(a, k, s);
}

}  // namespace aknn
