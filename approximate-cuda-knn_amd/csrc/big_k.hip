// Initial leaf merge and propagation rounds for 64 < k <= 256.
//
// The register and LDS layouts of the other kernels hold one list entry per
// lane (k <= 64).  Here every point's list lives in LDS as packed (key, id)
// pairs (topk.hpp's order), and new candidates are merged by a sort: the
// point's k pairs and a chunk of up to 512 - k candidates are sorted
// together (bitonic, one wave), and the first k distinct pairs are kept.  A
// candidate already in the list carries the same exact key as its stored
// pair (the same chain computed it), so the two copies are adjacent after
// the sort and the second is dropped; stub pairs (id -1) are all alike and
// all kept.  The first k of (list + all candidates) are the first k of (list
// + each chunk's candidates) taken chunk by chunk, so the result is the
// selection of the other kernels: exact keys, ties by lower id.
//
// Reference: findInitialStateOfAproximated{Euclidean,Taxicab,Cosine}KNN and
// the propagation kernels (euclideandistancematrixgpu.cu:107-1031,
// :1033-1935), whose scratch is sized for any k (dimIdx = min(k^2 + k, N),
// .cu:2768, :2797-2805).
#include "device_util.hpp"
#include "kernels.hpp"
#include "topk.hpp"
#include "wave_ops.hpp"

namespace aknn {

namespace {

constexpr int kPool = 512;  // pairs per wave: the list and a chunk of candidates
constexpr int kWaves = 4;   // waves per block, each on its own point

__device__ __forceinline__ uint64_t stub_pair() { return pack_pair(kStubDist, order_id(kStubId)); }

// The 64-bit value of the lane below (lane 0: its own).
__device__ __forceinline__ uint64_t shfl_up1_u64(uint64_t v) {
  const unsigned lo = __shfl_up(static_cast<unsigned>(v), 1), hi = __shfl_up(static_cast<unsigned>(v >> 32), 1);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Ascending bitonic sort of a[0, kPool) by one wave.
__device__ void pool_sort(uint64_t* a) {
  const int lane = threadIdx.x & 63;
  for (int size = 2; size <= kPool; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
      for (int t0 = 0; t0 < kPool / 2; t0 += 64) {
        const int t = t0 + lane;
        const int i = 2 * stride * (t / stride) + (t % stride), j = i + stride;
        const bool up = (i & size) == 0;
        const uint64_t x = a[i], y = a[j];
        if ((x > y) == up) {
          a[i] = y;
          a[j] = x;
        }
      }
      wave_sync();
    }
  }
}

// After pool_sort: the first k distinct pairs of a, in order, into a[0, k)
// (an entry equal to its predecessor is a second copy of one candidate,
// dropped unless it is a stub; ~0 marks unused slots).  Entries move only
// to lower or equal positions, after every lane has read its own.
__device__ void pool_take(uint64_t* a, int k) {
  const int lane = threadIdx.x & 63;
  const uint64_t stub = stub_pair();
  int w = 0;
  uint64_t last = ~0ull;  // the entry before the current 64
  for (int base = 0; base < kPool && w < k; base += 64) {
    const uint64_t v = a[base + lane];
    uint64_t pv = shfl_up1_u64(v);
    if (lane == 0) pv = base == 0 ? ~0ull : last;
    const bool keep = v != ~0ull && (v != pv || v == stub);
    const uint64_t b = __ballot(keep);
    const int r = __popcll(b & ((1ull << lane) - 1ull));
    last = lane_u64(v, 63);
    wave_sync();
    if (keep && w + r < k) a[w + r] = v;
    w += __popcll(b);
    wave_sync();
  }
}

// Each point's list and one tree's keys for its start range (key table
// row off[i]: key(i, perm[rs[i] + r]) at r), merged: one wave per point.
__global__ __launch_bounds__(64 * kWaves) void big_leaf_merge_kernel(const float* __restrict__ keys,
                                                                     const int64_t* __restrict__ off,
                                                                     const int32_t* __restrict__ perm,
                                                                     const int32_t* __restrict__ rs,
                                                                     const int32_t* __restrict__ rc, int64_t n, int k,
                                                                     int32_t* __restrict__ gids,
                                                                     float* __restrict__ gvals) {
  __shared__ uint64_t pool[kWaves][kPool];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kWaves + wave;
  if (i >= n) return;
  uint64_t* a = pool[wave];
  const int64_t row = i * k;
  for (int j = lane; j < k; j += 64) a[j] = pack_pair(gvals[row + j], order_id(gids[row + j]));
  const int32_t r0 = rs[i], cnt = rc[i];
  const float* kt = keys + off[i];
  const int cap = kPool - k;
  for (int c0 = 0; c0 < cnt; c0 += cap) {
    for (int j = lane; j < cap; j += 64) {
      uint64_t v = ~0ull;
      if (c0 + j < cnt) {
        const int32_t c = perm[r0 + c0 + j];
        if (c != i) v = pack_pair(kt[c0 + j], order_id(c));
      }
      a[k + j] = v;
    }
    wave_sync();
    pool_sort(a);
    pool_take(a, k);
  }
  for (int j = lane; j < k; j += 64) {
    gids[row + j] = pair_id(a[j]);
    gvals[row + j] = pair_key(a[j]);
  }
}

struct BigPropArgs {
  const float* x;
  int d, ld;
  const float* norms;
  int64_t n;
  int k;
  const int32_t* ids_in;
  const float* vals_in;
  int32_t* ids_out;
  float* vals_out;
};

// One Jacobi round: each point's row merged with its neighbours' rows'
// members (two hops, .cu:1078-1085), one wave per point.  Every candidate
// other than i itself and stubs is keyed exactly; a candidate already in the
// row carries its stored key and is dropped as a second copy.  (The stored
// value is the key for every metric taken here; COSINE parity, whose stored
// values are flipped, is refused for k > 64.)
template <int METRIC>
__global__ __launch_bounds__(64 * kWaves) void big_propagate_kernel(BigPropArgs a) {
  __shared__ uint64_t pool[kWaves][kPool];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kWaves + wave;
  if (i >= a.n) return;
  const int k = a.k;
  uint64_t* p = pool[wave];
  const int64_t row = i * k;
  for (int j = lane; j < k; j += 64) p[j] = pack_pair(a.vals_in[row + j], order_id(a.ids_in[row + j]));
  const float* xi = a.x + i * a.ld;
  const float ni = METRIC == kCosine ? a.norms[i] : 1.0f;
  const int kk = k * k, cap = kPool - k;
  for (int f0 = 0; f0 < kk; f0 += cap) {
    for (int j = lane; j < cap; j += 64) {
      const int f = f0 + j;
      uint64_t v = ~0ull;
      if (f < kk) {
        const int hop = f / k, l = f - hop * k;
        const int32_t nb = a.ids_in[row + hop];
        const int32_t c = nb >= 0 ? a.ids_in[static_cast<int64_t>(nb) * k + l] : -1;
        if (c >= 0 && c != i) {
          const float key = finish_key<METRIC>(row_chain<METRIC>(xi, a.x + static_cast<int64_t>(c) * a.ld, a.d), ni,
                                               METRIC == kCosine ? a.norms[c] : 1.0f);
          v = pack_pair(key, order_id(c));
        }
      }
      p[k + j] = v;
    }
    wave_sync();
    pool_sort(p);
    pool_take(p, k);
  }
  for (int j = lane; j < k; j += 64) {
    a.ids_out[row + j] = pair_id(p[j]);
    a.vals_out[row + j] = pair_key(p[j]);
  }
}

int64_t big_blocks(int64_t n) { return (n + kWaves - 1) / kWaves; }

}  // namespace

void launch_big_leaf_merge(const float* keys, const int64_t* off, const int32_t* perm, const int32_t* rs,
                           const int32_t* rc, int64_t n, int k, int32_t* ids, float* vals, hipStream_t s) {
  if (k <= 64 || k > kBigKMax) throw HipError("large-k leaf merge takes 64 < k <= 256");
  hipLaunchKernelGGL(big_leaf_merge_kernel, dim3(static_cast<unsigned>(big_blocks(n))), dim3(64 * kWaves), 0, s, keys,
                     off, perm, rs, rc, n, k, ids, vals);
  AKNN_HIP_TRY(hipGetLastError());
}

void launch_big_propagate(int metric, const float* x, int d, int ld, const float* norms, int64_t n, int k,
                          const int32_t* ids_in, const float* vals_in, int32_t* ids_out, float* vals_out,
                          hipStream_t s) {
  if (k <= 64 || k > kBigKMax) throw HipError("large-k propagation takes 64 < k <= 256");
  if (metric == kCosine && norms == nullptr) throw HipError("large-k propagation: COSINE needs the norms");
  const BigPropArgs a{x, d, ld, norms, n, k, ids_in, vals_in, ids_out, vals_out};
  const dim3 g(static_cast<unsigned>(big_blocks(n))), b(64 * kWaves);
  switch (metric) {
    case kEuclidean:
      hipLaunchKernelGGL(big_propagate_kernel<kEuclidean>, g, b, 0, s, a);
      break;
    case kTaxicab:
      hipLaunchKernelGGL(big_propagate_kernel<kTaxicab>, g, b, 0, s, a);
      break;
    case kTaxicabL1:
      hipLaunchKernelGGL(big_propagate_kernel<kTaxicabL1>, g, b, 0, s, a);
      break;
    case kCosine:
      hipLaunchKernelGGL(big_propagate_kernel<kCosine>, g, b, 0, s, a);
      break;
    default:
      throw HipError("unknown distance");
  }
  AKNN_HIP_TRY(hipGetLastError());
}

}  // namespace aknn
