// Normalization kernels: restate normalizeDataStep1/2
// (euclideandistancematrixgpu.cu:19-42, :44-85) on a row-major [n][ld] matrix.
//
// Step 1 keeps the reference's per-dimension sequential FP32 sum: one lane per
// dimension walks all points in order (the rounding of a sequential sum depends
// on the order, so this chain cannot be split).  The chain's speed is set by
// how many rows are in flight, so fifteen waves stage rows through LDS for the
// one wave that adds.  Option mean_fp64 replaces the sequential chain by
// blocked FP64 sums.  Step 2 is two passes over the uploaded points, which
// stay as they are: the max-abs of the centred values over the dimensions
// below 256 (read only), then one pass per row that centres, scales and
// writes the normalized row, and forms its FP16 sketch from the values it
// wrote (the same arithmetic, element for element, as centring, scaling and
// sketching in separate passes).
#include "device_util.hpp"
#include "kernels.hpp"

namespace aknn {

namespace {

constexpr int64_t kMean64Rows = 2048;            // rows per FP64 partial sum (mean_fp64)
constexpr int kMeanLoaders = 15;                  // waves 1..15 load, wave 0 adds
constexpr int kMeanDims = 16;                     // dimensions per workgroup (one chain per lane)
constexpr int kMeanRowsPerLoad = 64 / kMeanDims;  // rows one 64-lane load covers (4)
constexpr int kMeanLoadsPerTile = 8;              // loads each loader wave issues per tile
constexpr int kMeanRows = kMeanLoaders * kMeanLoadsPerTile * kMeanRowsPerLoad;  // rows per tile (480)
constexpr int kMeanSlots = 4;                     // LDS ring of tiles
constexpr int kMeanAhead = 3;                     // tiles in flight ahead of the adder
constexpr int kMeanChunk = 24;                    // rows the adder holds in registers per buffer
static_assert(kMeanRows % (2 * kMeanChunk) == 0, "the adder's double buffer walks whole tiles");

// Waits until at most N vector-memory operations of this wave are pending
// (gfx9 encoding: vmcnt in bits 3:0 and 15:14; expcnt and lgkmcnt left open).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt is 6 bits");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Waits until this wave's LDS operations are done (vmcnt and expcnt left open).
__device__ __forceinline__ void wait_lgkmcnt0() { __builtin_amdgcn_s_waitcnt(15 | (3 << 14) | (7 << 4)); }

// Workgroup barrier without __syncthreads()'s fence, which would also wait for
// every load in flight; the callers wait for exactly what the other waves
// read next.  The empty asm statements keep the compiler from moving memory
// accesses across it.
__device__ __forceinline__ void ring_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

typedef __attribute__((address_space(3))) void* lds_void_ptr;
typedef __attribute__((address_space(1))) const void* global_void_ptr;

// One workgroup per 16 dimensions, so a 128-dimensional matrix spreads its
// chains over 8 CUs.  Lanes 0..15 of wave 0 hold the sequential sums (one
// lane per dimension) and add the rows of each tile in order: the
// reference's sequential FP32 sum, which cannot be split, so one dependent
// add per row is the bound.  Waves 1..15 keep it fed: each 64-lane
// global_load_lds moves 4 rows x 16 dimensions (lane i: row i / 16,
// dimension i % 16) straight from global memory into LDS as [row][16], three
// tiles ahead in a ring of four tile slots, so no registers bound the rows in
// flight.  Every load address is clamped into the matrix (rows past n and
// dimensions past d read valid data the adder ignores), so each wave issues
// exactly 8 loads per tile and waits with a fixed count.  The adder reads a
// full tile through two 16-row register buffers, loading one while it adds
// the other.
__global__ __launch_bounds__(1024) void mean_kernel(const float* __restrict__ x, int64_t n, int d, int ld,
                                                    float* __restrict__ mean) {
  __shared__ __attribute__((aligned(16))) float ring[kMeanSlots][kMeanRows][kMeanDims];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int d0 = blockIdx.x * kMeanDims;
  const int64_t ntiles = (n + kMeanRows - 1) / kMeanRows;
  // loader lane: row within a load, dimension (clamped)
  const int lrow = lane / kMeanDims;
  const int ldim = d0 + lane % kMeanDims < d ? d0 + lane % kMeanDims : d - 1;
  auto issue = [&](int64_t t) {
    float* slot = &ring[t % kMeanSlots][0][0];
#pragma unroll
    for (int j = 0; j < kMeanLoadsPerTile; ++j) {
      const int r = ((wave - 1) * kMeanLoadsPerTile + j) * kMeanRowsPerLoad;  // first row of this load
      int64_t row = t * kMeanRows + r + lrow;
      row = row < n ? row : n - 1;
      __builtin_amdgcn_global_load_lds((global_void_ptr)(x + row * ld + ldim),
                                       (lds_void_ptr)(slot + r * kMeanDims), 4, 0, 0);
    }
  };
  if (wave > 0) {
#pragma unroll
    for (int t = 0; t < kMeanAhead; ++t) issue(t);
    wait_vmcnt<(kMeanAhead - 1) * kMeanLoadsPerTile>();  // tile 0 has landed
  }
  ring_barrier();
  // no scalar load may still be pending in the loop: LDS waits could not then
  // count down in order and would all become waits for zero
  wait_lgkmcnt0();
  float m = 0.0f;
  const int dl = lane % kMeanDims;
  for (int64_t t = 0; t < ntiles; ++t) {
    if (wave > 0) {
      // slot (t + 3) % 4 held tile t - 1, consumed before the last barrier
      issue(t + kMeanAhead);
      wait_vmcnt<(kMeanAhead - 1) * kMeanLoadsPerTile>();  // tile t + 1 has landed
    } else {
      const float* col = &ring[t % kMeanSlots][0][dl];
      const int64_t left = n - t * kMeanRows;
      if (left >= kMeanRows) {
        float a[kMeanChunk], b[kMeanChunk];
#pragma unroll
        for (int i = 0; i < kMeanChunk; ++i) a[i] = col[i * kMeanDims];
        wait_lgkmcnt0();
        // Each phase issues the other buffer's reads, adds a buffer that has
        // already landed, then waits: the reads have had the adds' time to
        // land, and the next phase's adds need no wait.  sched_barrier keeps
        // the compiler from regrouping the phases.
#pragma unroll
        for (int c = 0; c < kMeanRows; c += 2 * kMeanChunk) {
#pragma unroll
          for (int i = 0; i < kMeanChunk; ++i) b[i] = col[(c + kMeanChunk + i) * kMeanDims];
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int i = 0; i < kMeanChunk; ++i) m = m + a[i];
          wait_lgkmcnt0();
          __builtin_amdgcn_sched_barrier(0);
          if (c + 2 * kMeanChunk < kMeanRows) {
#pragma unroll
            for (int i = 0; i < kMeanChunk; ++i) a[i] = col[(c + 2 * kMeanChunk + i) * kMeanDims];
          }
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int i = 0; i < kMeanChunk; ++i) m = m + b[i];
          wait_lgkmcnt0();
          __builtin_amdgcn_sched_barrier(0);
        }
      } else {
        const int rows = static_cast<int>(left);
        for (int r = 0; r < rows; ++r) m = m + col[r * kMeanDims];
      }
      wait_lgkmcnt0();  // the slot is free once its reads are done
    }
    ring_barrier();
  }
  if (wave > 0) wait_vmcnt<0>();  // no load may land after the workgroup ends
  if (wave == 0 && lane < kMeanDims && d0 + lane < d) mean[d0 + lane] = m / static_cast<float>(n);
}

// dm: dimensions that enter the max-abs and get scaled (the reference:
// min(D, 256), .cu:57-65, :76-84; option scale_all: D)
// Mean in FP64 (option mean_fp64, not the reference's): block b sums rows
// [b * kMean64Rows, (b+1) * kMean64Rows) of every dimension in row order in
// FP64; the block sums are then added in block order and the mean rounded
// once to FP32.  The order is fixed by n alone.
__global__ __launch_bounds__(256) void mean64_partial_kernel(const float* __restrict__ x, int64_t n, int d, int ld,
                                                            double* __restrict__ part) {
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kMean64Rows;
  const int64_t r1 = r0 + kMean64Rows < n ? r0 + kMean64Rows : n;
  for (int dd = threadIdx.x; dd < d; dd += 256) {
    double sum = 0.0;
    int64_t r = r0;
    for (; r + 8 <= r1; r += 8) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = x[(r + u) * ld + dd];
#pragma unroll
      for (int u = 0; u < 8; ++u) sum = sum + static_cast<double>(v[u]);
    }
    for (; r < r1; ++r) sum = sum + static_cast<double>(x[r * ld + dd]);
    part[static_cast<int64_t>(blockIdx.x) * d + dd] = sum;
  }
}

// One workgroup per dimension: the block sums are read into LDS 2048 at a
// time by all threads, and one thread adds them in block order (a dependent
// FP64 add per block instead of a dependent global load).
constexpr int kMean64Stage = 2048;
__global__ __launch_bounds__(256) void mean64_final_kernel(const double* __restrict__ part, int64_t blocks, int64_t n,
                                                          int d, float* __restrict__ mean) {
  __shared__ double buf[kMean64Stage];
  const int dd = blockIdx.x;
  double sum = 0.0;
  for (int64_t b0 = 0; b0 < blocks; b0 += kMean64Stage) {
    const int cnt = blocks - b0 < kMean64Stage ? static_cast<int>(blocks - b0) : kMean64Stage;
    for (int j = threadIdx.x; j < cnt; j += 256) buf[j] = part[(b0 + j) * d + dd];
    __syncthreads();
    if (threadIdx.x == 0)
      for (int j = 0; j < cnt; ++j) sum = sum + buf[j];
    __syncthreads();
  }
  if (threadIdx.x == 0) mean[dd] = static_cast<float>(sum / static_cast<double>(n));
}

// Max-abs of the centred values x - mean over dimensions below dm, 16 lanes
// per row reading 16-byte pieces (16 rows per block, grid-stride), into gmax
// (non-negative floats order like their bits).
__global__ __launch_bounds__(256) void maxabs_kernel(const float* __restrict__ x, int64_t n, int ld, int dm,
                                                     const float* __restrict__ mean, uint32_t* __restrict__ gmax) {
  __shared__ float red[256];
  const int sub = threadIdx.x & 15;
  float local = 0.0f;
  const int pieces = (dm + 3) / 4;  // ld % 4 == 0 and dm <= ld: a piece that starts below dm stays in the row
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * 16 + (threadIdx.x >> 4); i < n;
       i += static_cast<int64_t>(gridDim.x) * 16) {
    const float* xr = x + i * ld;
    for (int c = sub; c < pieces; c += 16) {
      const float4 v = *reinterpret_cast<const float4*>(xr + 4 * c);
      const float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (4 * c + e < dm) {
          const float a = __builtin_fabsf(vv[e] - mean[4 * c + e]);
          local = a > local ? a : local;
        }
      }
    }
  }
  red[threadIdx.x] = local;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s && red[threadIdx.x + s] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + s];
    __syncthreads();
  }
  // non-negative floats order like their bit patterns
  if (threadIdx.x == 0) atomicMax(gmax, __float_as_uint(red[0]));
}

// Row i of the normalized matrix, LPR lanes per row (64 / LPR rows per wave),
// 16-byte pieces: (src - mean) for dimensions below d, divided by the max-abs
// below dm; the row's padding copied.  SK: its FP16 sketch from the values
// written (sketch_element: rounding, flushes, exact error terms), eps[i] and
// nh[i] from the row's lanes' FP64 sums.
template <bool SK, int LPR>
__global__ __launch_bounds__(256) void apply_kernel(const float* __restrict__ src, float* __restrict__ x, int64_t n,
                                                    int d, int ld, int dm, const float* __restrict__ mean,
                                                    const uint32_t* __restrict__ gmax, int l1, _Float16* __restrict__ xh,
                                                    int ldh, float* __restrict__ eps, float* __restrict__ nh) {
  constexpr int kRows = 256 / LPR;  // rows per block
  const int sub = threadIdx.x % LPR;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kRows + threadIdx.x / LPR;
  const bool live = i < n;  // every lane stays for the row reductions
  const float g = __uint_as_float(*gmax);
  const float* sr = src + (live ? i : 0) * ld;
  float* xr = x + i * ld;
  double acc = 0.0, sq = 0.0;
  bool bad = false;
  const int span = (SK && ldh > ld ? ldh : ld) / 4;  // 16-byte pieces (ld and ldh are multiples of 4)
  for (int c = sub; c < span; c += LPR) {
    const int d0 = 4 * c;
    float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (d0 < ld) {
      const float4 s4 = *reinterpret_cast<const float4*>(sr + d0);
      v[0] = s4.x;
      v[1] = s4.y;
      v[2] = s4.z;
      v[3] = s4.w;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (d0 + e < d) {
          v[e] = v[e] - mean[d0 + e];
          if (d0 + e < dm) v[e] = v[e] / g;
        }
      }
      if (live) *reinterpret_cast<float4*>(xr + d0) = make_float4(v[0], v[1], v[2], v[3]);
    }
    if constexpr (SK) {
      if (d0 < ldh) {
        _Float16 h[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) h[e] = sketch_element(d0 + e < d ? v[e] : 0.0f, l1 != 0, acc, sq, bad);
        typedef _Float16 half4 __attribute__((ext_vector_type(4)));
        const half4 hv = {h[0], h[1], h[2], h[3]};
        if (live) *reinterpret_cast<half4*>(xh + i * ldh + d0) = hv;
      }
    }
  }
  if constexpr (SK) {
    for (int off = LPR / 2; off > 0; off >>= 1) {
      acc += __shfl_xor(acc, off);
      sq += __shfl_xor(sq, off);
    }
    // the row's flag: a lane mask of the wave, LPR bits per row
    const uint64_t bb = __ballot(bad);
    const int r0 = (threadIdx.x & 63) / LPR * LPR;
    const uint64_t rm = LPR == 64 ? ~0ull : ((1ull << LPR) - 1) << r0;
    bad = (bb & rm) != 0;
    if (live && sub == 0) {
      eps[i] = sketch_radius(acc, l1 != 0, bad);
      nh[i] = static_cast<float>(sq);
    }
  }
}

__global__ __launch_bounds__(256) void norms_kernel(const float* __restrict__ x, int64_t n, int d, int ld,
                                                    float* __restrict__ norms) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i < n) norms[i] = norm_chain(x + i * ld, d);
}

int grid_for(int64_t work, int64_t per_block, int cap) {
  int64_t g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return static_cast<int>(g > cap ? cap : g);
}

}  // namespace

void launch_normalize(const float* src, float* x, int64_t n, int d, int ld, float* mean, uint32_t* gmax,
                      bool scale_all, double* mean64_part, const SketchOut* sk, hipStream_t s) {
  const int dm = scale_all ? d : (d < 256 ? d : 256);
  AKNN_HIP_TRY(hipMemsetAsync(gmax, 0, sizeof(uint32_t), s));
  if (mean64_part != nullptr) {
    const int64_t blocks = (n + kMean64Rows - 1) / kMean64Rows;
    hipLaunchKernelGGL(mean64_partial_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, src, n, d, ld,
                       mean64_part);
    hipLaunchKernelGGL(mean64_final_kernel, dim3(d), dim3(256), 0, s, mean64_part, blocks, n, d, mean);
  } else {
    hipLaunchKernelGGL(mean_kernel, dim3((d + kMeanDims - 1) / kMeanDims), dim3(1024), 0, s, src, n, d, ld, mean);
  }
  AKNN_HIP_TRY(hipGetLastError());
  if (ld % 4 != 0) throw HipError("normalize: row stride must be a multiple of 4");
  hipLaunchKernelGGL(maxabs_kernel, dim3(grid_for(n, 16, 8192)), dim3(256), 0, s, src, n, ld, dm, mean, gmax);
  AKNN_HIP_TRY(hipGetLastError());
  // 16 lanes per row for rows of up to 64 pieces (16 rows per block), else 64
  const bool narrow = ld <= 256 && (sk == nullptr || sk->ldh <= 256);
  const int per_block = narrow ? 256 / 16 : 256 / 64;
  const dim3 rows(static_cast<unsigned>((n + per_block - 1) / per_block));
  if (sk != nullptr) {
    if (sk->ldh % 8 != 0 || sk->ldh < d) throw HipError("sketch: row stride must be a multiple of 8 covering d");
    if (narrow)
      hipLaunchKernelGGL((apply_kernel<true, 16>), rows, dim3(256), 0, s, src, x, n, d, ld, dm, mean, gmax,
                         sk->l1 ? 1 : 0, sk->xh, sk->ldh, sk->eps, sk->nh);
    else
      hipLaunchKernelGGL((apply_kernel<true, 64>), rows, dim3(256), 0, s, src, x, n, d, ld, dm, mean, gmax,
                         sk->l1 ? 1 : 0, sk->xh, sk->ldh, sk->eps, sk->nh);
  } else {
    if (narrow)
      hipLaunchKernelGGL((apply_kernel<false, 16>), rows, dim3(256), 0, s, src, x, n, d, ld, dm, mean, gmax, 0,
                         nullptr, 0, nullptr, nullptr);
    else
      hipLaunchKernelGGL((apply_kernel<false, 64>), rows, dim3(256), 0, s, src, x, n, d, ld, dm, mean, gmax, 0,
                         nullptr, 0, nullptr, nullptr);
  }
  AKNN_HIP_TRY(hipGetLastError());
}

void launch_norms(const float* x, int64_t n, int d, int ld, float* norms, hipStream_t s) {
  hipLaunchKernelGGL(norms_kernel, dim3(grid_for(n, 256, 1 << 30)), dim3(256), 0, s, x, n, d, ld, norms);
  AKNN_HIP_TRY(hipGetLastError());
}

}  // namespace aknn
