// recall@k of the current graph without any device allocation.
//
// The allocating form (Engine::recall) takes five device buffers from
// hipMalloc per call and frees them on return: the query indices, the exact
// k-NN ids and FP64 distances [nq][k], each query's k-th exact distance and
// the hit counter, plus the FP64 squared norms of all points for COSINE.
// Here the same kernels take those buffers from one static device array,
// `scratch`, of kScratchBytes per GPU, that stays with the process.  Callers
// that may not allocate (tight serving loops, harnesses that forbid hipMalloc)
// get the same result bit for bit.
//
// The array's host-side name is not a device address: its address on each
// device comes from hipGetSymbolAddress, taken once per device with that
// device current, and only that address is handed to kernels and copies.
// Every call on one GPU shares that GPU's array and leaves its contents
// behind; the hit counter is cleared by a kernel on every call, and every
// other region is fully rewritten before it is read.  Calls that overlap on
// one GPU (two host threads, two handles on the same device) race on the
// array and must be serialised by the caller.
#include <mutex>

#include "engine.hpp"
#include "kernels.hpp"

namespace aknn {

constexpr size_t kScratchBytes = size_t{48} << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// This device's copy of `scratch` (the device is current).
unsigned char* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) throw HipError("no static scratch for device " + std::to_string(device));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    AKNN_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw HipError("hipGetSymbolAddress returned no address for the static scratch");
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  }
  return g_scratch_ptr[device];
}

__global__ void clear_counter_kernel(unsigned long long* c) {
  if (threadIdx.x == 0) c[0] = 0;
}

constexpr size_t round256(size_t v) { return (v + 255) / 256 * 256; }

}  // namespace

size_t recall_scratch_bytes(int64_t n, int k, int64_t nq, bool cosine) {
  const size_t q = static_cast<size_t>(nq);
  size_t b = round256(q * sizeof(int64_t));                               // query indices
  b += round256(q * k * sizeof(int32_t));                                 // exact ids
  b += round256(q * k * sizeof(double));                                  // exact distances
  b += round256(q * sizeof(double));                                      // k-th exact distance
  b += 256;                                                               // hit counter
  if (cosine) b += round256(static_cast<size_t>(n) * sizeof(double));   // FP64 squared norms
  return b;
}

double recall_no_alloc(Engine& e, const int64_t* queries, int64_t nq) {
  const int64_t n = e.n();
  const int k = e.params().k;
  const bool cosine = e.params().metric == kCosine;
  if (nq < 0) throw std::invalid_argument("negative query count");
  for (int64_t q = 0; q < nq; ++q)
    if (queries[q] < 0 || queries[q] >= n) throw std::invalid_argument("query index out of range");
  if (nq == 0) return 0.0;
  const size_t need = recall_scratch_bytes(n, k, nq, cosine);
  if (need > kScratchBytes)
    throw std::invalid_argument("recall without allocation needs " + std::to_string(need) +
                                " bytes of scratch for these queries, the static scratch holds " +
                                std::to_string(kScratchBytes) + " (use fewer queries)");
  AKNN_HIP_TRY(hipSetDevice(e.device()));
  unsigned char* base = scratch_on(e.device());
  hipStream_t s = e.stream();
  size_t o = 0;
  int64_t* dq = reinterpret_cast<int64_t*>(base + o);
  o += round256(nq * sizeof(int64_t));
  int32_t* di = reinterpret_cast<int32_t*>(base + o);
  o += round256(static_cast<size_t>(nq) * k * sizeof(int32_t));
  double* dd = reinterpret_cast<double*>(base + o);
  o += round256(static_cast<size_t>(nq) * k * sizeof(double));
  double* kth = reinterpret_cast<double*>(base + o);
  o += round256(nq * sizeof(double));
  unsigned long long* hits = reinterpret_cast<unsigned long long*>(base + o);
  o += 256;
  double* sq = nullptr;
  if (cosine) {
    sq = reinterpret_cast<double*>(base + o);
    launch_sq64(e.points(), n, e.d(), e.ld(), sq, s);
  }
  AKNN_HIP_TRY(hipMemcpyAsync(dq, queries, sizeof(int64_t) * nq, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(clear_counter_kernel, dim3(1), dim3(64), 0, s, hits);
  AKNN_HIP_TRY(hipGetLastError());
  launch_exact_knn(e.params().metric, e.points(), n, e.d(), e.ld(), sq, dq, nq, k, di, dd, s);
  // the k-th exact distance of query q is dd[q * k + k - 1]
  AKNN_HIP_TRY(hipMemcpy2DAsync(kth, sizeof(double), dd + (k - 1), sizeof(double) * k, sizeof(double), nq,
                                hipMemcpyDeviceToDevice, s));
  launch_recall(e.params().metric, e.points(), e.d(), e.ld(), sq, dq, nq, kth, e.graph_ids(), k, hits, s);
  unsigned long long h = 0;
  AKNN_HIP_TRY(hipMemcpyAsync(&h, hits, sizeof(h), hipMemcpyDeviceToHost, s));
  AKNN_HIP_TRY(hipStreamSynchronize(s));
  return static_cast<double>(h) / (static_cast<double>(nq) * k);
}

}  // namespace aknn
