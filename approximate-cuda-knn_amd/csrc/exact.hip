// Exact k-NN in FP64 of query points against all points, and the recall of a
// graph against it: the reference that recall@k of the build is measured on.
//
// Distances are taken on the normalized points the build sees, in FP64, with
// one fixed evaluation order: dimensions in order, EUCLIDEAN the squared sum
// s = s + t*t (t = a - b), TAXICAB s = s + |t|, COSINE 1 - ab / (sqrt(aa)
// sqrt(bb)) with the three sums in order (multiply and add rounded
// separately, no contraction).  Every distance this file computes comes from
// that one function, so a neighbour whose distance ties the k-th exact
// distance counts as a hit, consistently.  The top-k is by (distance, id);
// the query itself is excluded.
//
// MI355X mapping: a block holds 16 (or fewer, for wide rows) query rows in
// LDS as FP64; its four waves each stream a quarter of the points, one point
// per lane, FP32 rows read in 16-byte pieces and widened, and keep one sorted
// list per query spread over the lanes.  The four partial lists are merged in
// LDS at the end.
#include <climits>

#include "device_util.hpp"
#include "kernels.hpp"

namespace aknn {

namespace {

constexpr int kExactWaves = 4;

__device__ __forceinline__ bool dkey_less(double ka, int32_t ia, double kb, int32_t ib) {
  return ka < kb || (ka == kb && ia < ib);
}

// Sorted list of (key, id) held by lanes [0, k) of the wave; inserts an
// entry strictly below the last one (which is dropped).
__device__ __forceinline__ void dlist_insert(double& key, int32_t& id, int k, double nk, int32_t nid) {
  const int lane = lane_id();
  const bool smaller = lane < k && dkey_less(key, id, nk, nid);
  const int pos = __popcll(__ballot(smaller));
  const double pk = __shfl_up(key, 1);
  const int32_t pid = __shfl_up(id, 1);
  if (lane < k && lane > pos) {
    key = pk;
    id = pid;
  }
  if (lane == pos) {
    key = nk;
    id = nid;
  }
}

// Squared norms in FP64, dimensions in order (COSINE).
__global__ __launch_bounds__(256) void sq64_kernel(const float* __restrict__ x, int64_t n, int d, int ld,
                                                   double* __restrict__ sq) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int dd = 0; dd < d; ++dd) {
    const double v = x[i * ld + dd];
    s = s + v * v;
  }
  sq[i] = s;
}

template <int METRIC>
__device__ __forceinline__ double finish64(double s, double ab, double aa, double bb) {
  if constexpr (METRIC == kCosine) {
    (void)s;
    return 1.0 - ab / (sqrt(aa) * sqrt(bb));
  } else {
    (void)ab;
    (void)aa;
    (void)bb;
    return s;
  }
}

// One pair's FP64 distance, both rows in global memory (FP32).
template <int METRIC>
__device__ __forceinline__ double dist64(const float* __restrict__ a, const float* __restrict__ b, int d, double aa,
                                         double bb) {
  double s = 0.0;
  for (int dd = 0; dd < d; ++dd) {
    const double p = a[dd], q = b[dd];
    if constexpr (METRIC == kEuclidean) {
      const double t = p - q;
      s = s + t * t;
    } else if constexpr (METRIC == kTaxicab) {
      s = s + __builtin_fabs(p - q);
    } else {
      s = s + p * q;
    }
  }
  return METRIC == kCosine ? finish64<METRIC>(0.0, s, aa, bb) : s;
}

template <int METRIC, int QB>
__global__ __launch_bounds__(256) void exact_knn_kernel(const float* __restrict__ x, int64_t n, int d, int ld,
                                                        const double* __restrict__ sq,
                                                        const int64_t* __restrict__ queries, int64_t nq, int kq,
                                                        int32_t* __restrict__ out_ids, double* __restrict__ out_d) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  double* qrow = reinterpret_cast<double*>(smem);                  // [QB][ld]
  double* mkey = qrow + static_cast<size_t>(QB) * ld;              // [waves][QB][64]
  int32_t* mid = reinterpret_cast<int32_t*>(mkey + kExactWaves * QB * 64);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * QB;
  int64_t qid[QB];
  double qaa[QB];
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    qid[q] = q0 + q < nq ? queries[q0 + q] : -1;
    qaa[q] = METRIC == kCosine && qid[q] >= 0 ? sq[qid[q]] : 1.0;
  }
  for (int e = threadIdx.x; e < QB * ld; e += blockDim.x) {
    const int q = e / ld, dd = e - q * ld;
    const int64_t qi = q0 + q < nq ? queries[q0 + q] : -1;
    qrow[e] = qi >= 0 && dd < d ? static_cast<double>(x[qi * ld + dd]) : 0.0;
  }
  __syncthreads();

  double key[QB];
  int32_t kid[QB];
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    key[q] = __builtin_inf();
    kid[q] = INT_MAX;
  }
  const int64_t per = (n + kExactWaves - 1) / kExactWaves;
  const int64_t p0 = wave * per, p1 = p0 + per < n ? p0 + per : n;
  for (int64_t base = p0; base < p1; base += 64) {
    const int64_t p = base + lane;
    const bool valid = p < p1;
    double acc[QB];
#pragma unroll
    for (int q = 0; q < QB; ++q) acc[q] = 0.0;
    const float* xp = x + (valid ? p : 0) * ld;
    for (int dd = 0; dd < d; dd += 4) {
      const float4 v4 = *reinterpret_cast<const float4*>(xp + dd);
      const double pv[4] = {v4.x, v4.y, v4.z, v4.w};
      const int nd = d - dd < 4 ? d - dd : 4;
#pragma unroll
      for (int q = 0; q < QB; ++q) {
        const double* qr = qrow + q * ld + dd;
        for (int j = 0; j < nd; ++j) {
          if constexpr (METRIC == kEuclidean) {
            const double t = qr[j] - pv[j];
            acc[q] = acc[q] + t * t;
          } else if constexpr (METRIC == kTaxicab) {
            acc[q] = acc[q] + __builtin_fabs(qr[j] - pv[j]);
          } else {
            acc[q] = acc[q] + qr[j] * pv[j];
          }
        }
      }
    }
    const double pbb = METRIC == kCosine && valid ? sq[p] : 1.0;
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const double dv = finish64<METRIC>(acc[q], acc[q], qaa[q], pbb);
      const double thk = __shfl(key[q], kq - 1);
      const int32_t thi = __shfl(kid[q], kq - 1);
      const int32_t pid = static_cast<int32_t>(p);
      uint64_t surv = __ballot(valid && p != qid[q] && qid[q] >= 0 && dkey_less(dv, pid, thk, thi));
      while (surv) {
        const int src = __ffsll(static_cast<long long>(surv)) - 1;
        surv &= surv - 1;
        const double nk = __shfl(dv, src);
        const int32_t nid = __shfl(pid, src);
        if (dkey_less(nk, nid, __shfl(key[q], kq - 1), __shfl(kid[q], kq - 1)))
          dlist_insert(key[q], kid[q], kq, nk, nid);
      }
    }
  }
  // merge the waves' lists: each wave publishes, wave 0 inserts the others'
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    mkey[(wave * QB + q) * 64 + lane] = key[q];
    mid[(wave * QB + q) * 64 + lane] = kid[q];
  }
  __syncthreads();
  if (wave != 0) return;
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    for (int w = 1; w < kExactWaves; ++w) {
      const double ok = lane < kq ? mkey[(w * QB + q) * 64 + lane] : __builtin_inf();
      const int32_t oi = lane < kq ? mid[(w * QB + q) * 64 + lane] : INT_MAX;
      uint64_t surv = __ballot(lane < kq && dkey_less(ok, oi, __shfl(key[q], kq - 1), __shfl(kid[q], kq - 1)));
      while (surv) {
        const int src = __ffsll(static_cast<long long>(surv)) - 1;
        surv &= surv - 1;
        const double nk = __shfl(ok, src);
        const int32_t nid = __shfl(oi, src);
        if (dkey_less(nk, nid, __shfl(key[q], kq - 1), __shfl(kid[q], kq - 1)))
          dlist_insert(key[q], kid[q], kq, nk, nid);
      }
    }
    if (q0 + q < nq && lane < kq) {
      out_ids[(q0 + q) * kq + lane] = kid[q] == INT_MAX ? -1 : kid[q];
      out_d[(q0 + q) * kq + lane] = key[q];
    }
  }
}

// hits[0] += number of graph neighbours of each query within its k-th exact
// distance (one wave per query, one lane per neighbour)
template <int METRIC>
__global__ __launch_bounds__(64) void recall_kernel(const float* __restrict__ x, int d, int ld,
                                                    const double* __restrict__ sq,
                                                    const int64_t* __restrict__ queries, int64_t nq,
                                                    const double* __restrict__ kth, const int32_t* __restrict__ ids,
                                                    int k, unsigned long long* __restrict__ hits) {
  const int64_t qi = blockIdx.x;
  if (qi >= nq) return;
  const int lane = threadIdx.x;
  const int64_t q = queries[qi];
  bool hit = false;
  if (lane < k) {
    const int32_t c = ids[q * k + lane];
    if (c >= 0 && c != q) {
      const double dv = dist64<METRIC>(x + q * ld, x + static_cast<int64_t>(c) * ld, d,
                                       METRIC == kCosine ? sq[q] : 1.0, METRIC == kCosine ? sq[c] : 1.0);
      hit = dv <= kth[qi];
    }
  }
  const int h = __popcll(__ballot(hit));
  if (lane == 0 && h > 0) atomicAdd(hits, static_cast<unsigned long long>(h));
}

template <int METRIC, int QB>
void launch_q(const float* x, int64_t n, int d, int ld, const double* sq, const int64_t* queries, int64_t nq, int kq,
              int32_t* ids, double* dist, hipStream_t s) {
  const size_t shmem = (static_cast<size_t>(QB) * ld + kExactWaves * QB * 64) * sizeof(double) +
                       kExactWaves * QB * 64 * sizeof(int32_t);
  if (shmem > 160 * 1024) throw HipError("exact k-NN: rows too wide");
  const int64_t blocks = (nq + QB - 1) / QB;
  hipLaunchKernelGGL((exact_knn_kernel<METRIC, QB>), dim3(static_cast<unsigned>(blocks)), dim3(64 * kExactWaves),
                     shmem, s, x, n, d, ld, sq, queries, nq, kq, ids, dist);
  AKNN_HIP_TRY(hipGetLastError());
}

template <int METRIC>
void launch_m(const float* x, int64_t n, int d, int ld, const double* sq, const int64_t* queries, int64_t nq, int kq,
              int32_t* ids, double* dist, hipStream_t s) {
  if (ld <= 512)
    launch_q<METRIC, 16>(x, n, d, ld, sq, queries, nq, kq, ids, dist, s);
  else if (ld <= 1024)
    launch_q<METRIC, 8>(x, n, d, ld, sq, queries, nq, kq, ids, dist, s);
  else
    launch_q<METRIC, 4>(x, n, d, ld, sq, queries, nq, kq, ids, dist, s);
}

}  // namespace

void launch_sq64(const float* x, int64_t n, int d, int ld, double* sq, hipStream_t s) {
  hipLaunchKernelGGL(sq64_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, x, n, d, ld, sq);
  AKNN_HIP_TRY(hipGetLastError());
}

void launch_exact_knn(int metric, const float* x, int64_t n, int d, int ld, const double* sq, const int64_t* queries,
                      int64_t nq, int kq, int32_t* ids, double* dist, hipStream_t s) {
  if (kq < 1 || kq > 64) throw HipError("exact k-NN supports 1 <= k <= 64");
  if (nq <= 0) return;
  switch (metric) {
    case kEuclidean: launch_m<kEuclidean>(x, n, d, ld, sq, queries, nq, kq, ids, dist, s); break;
    case kTaxicab: launch_m<kTaxicab>(x, n, d, ld, sq, queries, nq, kq, ids, dist, s); break;
    case kCosine: launch_m<kCosine>(x, n, d, ld, sq, queries, nq, kq, ids, dist, s); break;
    default: throw HipError("unknown distance");
  }
}

void launch_recall(int metric, const float* x, int d, int ld, const double* sq, const int64_t* queries, int64_t nq,
                   const double* kth, const int32_t* ids, int k, unsigned long long* hits, hipStream_t s) {
  if (nq <= 0) return;
  if (k > 64) throw HipError("recall supports k <= 64");
  switch (metric) {
    case kEuclidean:
      hipLaunchKernelGGL(recall_kernel<kEuclidean>, dim3(static_cast<unsigned>(nq)), dim3(64), 0, s, x, d, ld, sq,
                         queries, nq, kth, ids, k, hits);
      break;
    case kTaxicab:
      hipLaunchKernelGGL(recall_kernel<kTaxicab>, dim3(static_cast<unsigned>(nq)), dim3(64), 0, s, x, d, ld, sq,
                         queries, nq, kth, ids, k, hits);
      break;
    case kCosine:
      hipLaunchKernelGGL(recall_kernel<kCosine>, dim3(static_cast<unsigned>(nq)), dim3(64), 0, s, x, d, ld, sq,
                         queries, nq, kth, ids, k, hits);
      break;
    default:
      throw HipError("unknown distance");
  }
  AKNN_HIP_TRY(hipGetLastError());
}

}  // namespace aknn
