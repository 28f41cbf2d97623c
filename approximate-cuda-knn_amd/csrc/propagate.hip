// One neighbour-propagation round (Jacobi: every read uses the previous graph).
//
// Reference: propagate{Euclidean,Taxicab,Cosine}Kernel
// (euclideandistancematrixgpu.cu:1033-1326, :1328-1621, :1623-1935), host loop
// .cu:1937-1977.  For point i the new candidates are the neighbours of its
// neighbours minus i minus its own neighbours (.cu:1078-1100, an N-byte bitmap
// per block there); if there are none the row is copied (.cu:1117); otherwise
// the k carried entries (keyed by their stored values) and the candidates
// (keyed by their distance) compete, and the first k by (key, id) are stored
// (COSINE parity stores 1 - key, .cu:1911).
//
// MI355X mapping: one wave per point.  The k*k two-hop ids are gathered with
// one load per lane-slot, deduplicated in an LDS hash set that was seeded with
// i and i's own row, and compacted in LDS.  Lanes then take one candidate each:
// the candidate's row is streamed from HBM with 16-byte loads while x_i is read
// from LDS by broadcast, giving the exact FP32 chain.  The running top-k lives
// one entry per lane; candidates below the current k-th entry (few after the
// first round) are inserted with ballot/shuffle steps.
#include <climits>

#include "device_util.hpp"
#include "kernels.hpp"

namespace aknn {

namespace {

constexpr int kWavesPerBlock = 4;

__device__ __forceinline__ bool hash_insert(int32_t* hs, uint32_t mask, int shift, int32_t key) {
  uint32_t h = (static_cast<uint32_t>(key) * 2654435761u) >> shift;
  for (;;) {
    const int32_t old = atomicCAS(&hs[h], -1, key);
    if (old == -1) return true;
    if (old == key) return false;
    h = (h + 1) & mask;
  }
}

template <int METRIC, bool FLIP, int NV>
__global__ __launch_bounds__(256) void propagate_kernel(const float* __restrict__ x, int d, int ld,
                                                        const float* __restrict__ norms,
                                                        const int32_t* __restrict__ order, int64_t n, int k,
                                                        int hs_log2, int cand_cap, int xi_stride,
                                                        const int32_t* __restrict__ ids_in,
                                                        const float* __restrict__ vals_in, int32_t* ids_out,
                                                        float* __restrict__ vals_out,
                                                        unsigned long long* __restrict__ ncand,
                                                        const int32_t* ids_prev,
                                                        const uint8_t* __restrict__ changed_prev,
                                                        uint8_t* __restrict__ changed_out) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int hs_size = 1 << hs_log2;
  const size_t per_wave_words = static_cast<size_t>(hs_size) + cand_cap + xi_stride + 4;
  float* xi = reinterpret_cast<float*>(smem) + per_wave_words * wave;  // 16-byte aligned (xi_stride % 4 == 0)
  int32_t* hs = reinterpret_cast<int32_t*>(xi + xi_stride);
  int32_t* cand = hs + hs_size;
  int32_t* cnt = cand + cand_cap;

  // Points are taken in `order` (the first tree's leaf order): waves running at
  // the same time then own nearby points whose candidate rows overlap, and
  // each XCD walks a contiguous part of that order, so its L2 sees the reuse.
  const int64_t pidx = xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + wave;
  if (pidx >= n) return;
  const int64_t i = order != nullptr ? order[pidx] : pidx;

  float key = __builtin_inff();
  int32_t id = INT_MAX;
  if (lane < k) {
    key = vals_in[i * k + lane];
    id = ids_in[i * k + lane];
  }
  // Neighbours whose whole row was already a candidate set of i last round
  // need no expansion: j was in i's previous row and j's row did not change
  // in the previous round.  Every such candidate was then in last round's
  // pool of i, and a candidate that lost there has k better entries in i's
  // current row, so it cannot enter now (ids_prev is the buffer this round
  // writes; row i of it is read here before this wave overwrites it).
  bool expand = true;
  if (ids_prev != nullptr) {
    const int32_t pid = lane < k ? ids_prev[i * k + lane] : INT_MAX;
    bool was_prev = false;
    for (int l = 0; l < k; ++l) was_prev |= __shfl(pid, l) == id;
    if (lane < k && id >= 0 && was_prev && changed_prev[id] == 0) expand = false;
  }
  for (int s = lane; s < hs_size; s += 64) hs[s] = -1;
  for (int dd = lane; dd < d; dd += 64) xi[dd] = x[i * ld + dd];
  if (lane == 0) *cnt = 0;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  const uint32_t mask = static_cast<uint32_t>(hs_size - 1);
  const int shift = 32 - hs_log2;
  if (lane == 0) hash_insert(hs, mask, shift, static_cast<int32_t>(i));
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (lane < k && id >= 0) hash_insert(hs, mask, shift, id);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  // two-hop ids (.cu:1078-1085)
  const int kk = k * k;
  for (int f = lane; f < kk + (64 - kk % 64) % 64; f += 64) {
    const int hi = f / k, lo = f - (f / k) * k;
    const int32_t nb = __shfl(id, hi < 64 ? hi : 0);
    const bool ex = __shfl(static_cast<int>(expand), hi < 64 ? hi : 0) != 0;
    if (f < kk && nb >= 0 && ex) {
      const int32_t c = ids_in[static_cast<int64_t>(nb) * k + lo];
      if (c >= 0 && hash_insert(hs, mask, shift, c)) {
        const int pos = atomicAdd(cnt, 1);
        cand[pos] = c;
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const int m = *cnt;
  if (ncand != nullptr && lane == 0) atomicAdd(ncand, static_cast<unsigned long long>(m));

  bool changed = false;
  if (m > 0) {
    float thk = __shfl(key, k - 1);
    int32_t thi = __shfl(id, k - 1);
    const float ni = METRIC == kCosine ? norms[i] : 1.0f;
    for (int base = 0; base < m; base += 64) {
      const int ci = base + lane;
      const int32_t c = ci < m ? cand[ci] : -1;
      float dk = __builtin_inff();
      if (c >= 0) {
        const float* xc = x + static_cast<int64_t>(c) * ld;
        float t;
        if constexpr (NV > 0)
          t = row_chain_regs<METRIC, NV>(xi, xc);
        else
          t = row_chain<METRIC>(xi, xc, d);
        dk = finish_key<METRIC>(t, ni, METRIC == kCosine ? norms[c] : 1.0f);
      }
      uint64_t surv = __ballot(c >= 0 && key_less(dk, c, thk, thi));
      while (surv) {
        const int src = __ffsll(static_cast<long long>(surv)) - 1;
        surv &= surv - 1;
        const float nk = __shfl(dk, src);
        const int32_t nid = __shfl(c, src);
        if (key_less(nk, nid, thk, thi)) {
          changed = true;
          wave_list_insert(key, id, k, nk, nid);
          thk = __shfl(key, k - 1);
          thi = __shfl(id, k - 1);
        }
      }
    }
    if constexpr (FLIP) {
      if (lane < k) key = static_cast<float>(1.0 - static_cast<double>(key));
      wave_sort(key, id);
      changed = true;
    }
  }
  if (changed_out != nullptr && lane == 0) changed_out[i] = changed ? 1 : 0;
  if (lane < k) {
    ids_out[i * k + lane] = id;
    vals_out[i * k + lane] = key;
  }
}

template <int METRIC, bool FLIP, int NV>
void launch_w(const float* x, int d, int ld, const float* norms, const int32_t* order, int64_t n, int k,
              const int32_t* ids_in, const float* vals_in, int32_t* ids_out, float* vals_out,
              unsigned long long* ncand, const PropagateHistory& hist, hipStream_t s) {
  int hs_log2 = 7;
  while ((1 << hs_log2) < 2 * (k * k + k + 1)) ++hs_log2;
  const int cand_cap = k * k;
  const int xi_stride = (d + 3) / 4 * 4;
  const size_t per_wave = (static_cast<size_t>(1 << hs_log2) + cand_cap + xi_stride + 4) * 4;
  const size_t shmem = per_wave * kWavesPerBlock;
  if (shmem > 160 * 1024) throw HipError("propagation: k or d too large for LDS");
  const int64_t blocks = xcd_grid((n + kWavesPerBlock - 1) / kWavesPerBlock);
  hipLaunchKernelGGL((propagate_kernel<METRIC, FLIP, NV>), dim3(static_cast<unsigned>(blocks)), dim3(256), shmem, s, x, d,
                     ld, norms, order, n, k, hs_log2, cand_cap, xi_stride, ids_in, vals_in, ids_out, vals_out, ncand,
                     FLIP ? nullptr : hist.ids_prev, hist.changed_prev, hist.changed_out);
  AKNN_HIP_TRY(hipGetLastError());
}

// Candidate rows are streamed (width class 0) for every width: the
// register-row instantiations measured slower here (fewer waves in flight at
// 196 VGPRs outweigh the deeper per-lane prefetch).
template <int METRIC, bool FLIP>
void launch_t(const float* x, int d, int ld, const float* norms, const int32_t* order, int64_t n, int k,
              const int32_t* ids_in, const float* vals_in, int32_t* ids_out, float* vals_out,
              unsigned long long* ncand, const PropagateHistory& hist, hipStream_t s) {
  launch_w<METRIC, FLIP, 0>(x, d, ld, norms, order, n, k, ids_in, vals_in, ids_out, vals_out, ncand, hist, s);
}

}  // namespace

void launch_propagate(int metric, bool flip, const float* x, int d, int ld, const float* norms,
                      const int32_t* order, int64_t n, int k, const int32_t* ids_in, const float* vals_in,
                      int32_t* ids_out, float* vals_out, unsigned long long* ncand, const PropagateHistory& hist,
                      hipStream_t s) {
  if (k > 64) throw HipError("propagation supports k <= 64");
  switch (metric) {
    case kEuclidean:
      launch_t<kEuclidean, false>(x, d, ld, norms, order, n, k, ids_in, vals_in, ids_out, vals_out, ncand, hist, s);
      break;
    case kTaxicab:
      launch_t<kTaxicab, false>(x, d, ld, norms, order, n, k, ids_in, vals_in, ids_out, vals_out, ncand, hist, s);
      break;
    case kCosine:
      if (flip)
        launch_t<kCosine, true>(x, d, ld, norms, order, n, k, ids_in, vals_in, ids_out, vals_out, ncand, hist, s);
      else
        launch_t<kCosine, false>(x, d, ld, norms, order, n, k, ids_in, vals_in, ids_out, vals_out, ncand, hist, s);
      break;
    default:
      throw HipError("unknown distance");
  }
}

}  // namespace aknn
