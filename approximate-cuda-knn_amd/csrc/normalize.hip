// Normalization kernels: restate normalizeDataStep1/2
// (euclideandistancematrixgpu.cu:19-42, :44-85) on a row-major [n][ld] matrix.
//
// Step 1 keeps the reference's per-dimension sequential FP32 sum: one lane per
// dimension walks all points in order (the rounding of a sequential sum depends
// on the order, so this chain cannot be split).  The chain's speed is set by
// how many rows are in flight, so fifteen waves stage rows through LDS for the
// one wave that adds.  Centring, the max-abs reduction over the dimensions
// below 256 and the scaling are plain parallel HBM passes.
#include "device_util.hpp"
#include "kernels.hpp"

namespace aknn {

namespace {

constexpr int kMeanRows = 256;  // rows per LDS tile (64 dimensions wide)
constexpr int kMeanPad = 4;     // floats of padding per LDS column (bank spread)
constexpr int kMeanLoaders = 15;
constexpr int kMeanPer = (kMeanRows + kMeanLoaders - 1) / kMeanLoaders;  // rows per loader lane per tile

// One workgroup per 64 dimensions.  Wave 0 holds the 64 sequential sums (one
// lane per dimension) and walks the rows of a tile in order; waves 1..15 load
// the rows.  The add chain is the reference's sequential order and cannot be
// split, so the aim is to keep it fed: each loader wave keeps four tiles of
// loads in flight in registers (slot u % 4 holds tile u) and writes the next
// tile into the free half of a two-tile LDS ring while wave 0 adds the other.
// The ring is column-major (one padded column of rows per dimension), so the
// adding lane reads four consecutive rows with one 16-byte LDS load.
__global__ __launch_bounds__(1024) void mean_kernel(const float* __restrict__ x, int64_t n, int d, int ld,
                                                    float* __restrict__ mean) {
  __shared__ __attribute__((aligned(16))) float ring[2][64][kMeanRows + kMeanPad];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int dd = blockIdx.x * 64 + lane;
  const int64_t ntiles = (n + kMeanRows - 1) / kMeanRows;
  float v[4][kMeanPer];
  auto issue = [&](float (&dst)[kMeanPer], int64_t t) {
#pragma unroll
    for (int j = 0; j < kMeanPer; ++j) {
      const int r = wave - 1 + kMeanLoaders * j;
      const int64_t row = t * kMeanRows + r;
      dst[j] = (t < ntiles && r < kMeanRows && row < n && dd < d) ? x[row * ld + dd] : 0.0f;
    }
  };
  auto store = [&](const float (&src)[kMeanPer], int buf) {
#pragma unroll
    for (int j = 0; j < kMeanPer; ++j) {
      const int r = wave - 1 + kMeanLoaders * j;
      if (r < kMeanRows) ring[buf][lane][r] = src[j];
    }
  };
  if (wave > 0) {
    issue(v[0], 0);
    issue(v[1], 1);
    issue(v[2], 2);
    issue(v[3], 3);
    store(v[0], 0);
    issue(v[0], 4);
  }
  __syncthreads();
  float m = 0.0f;
  for (int64_t t0 = 0; t0 < ntiles; t0 += 4) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int64_t t = t0 + s;
      if (t < ntiles) {
        if (wave > 0) {
          // tile t+1 from its slot into the free buffer, then tile t+5 into that slot
          if (t + 1 < ntiles) {
            store(v[(s + 1) & 3], static_cast<int>((t + 1) & 1));
            issue(v[(s + 1) & 3], t + 5);
          }
        } else {
          const int64_t left = n - t * kMeanRows;
          const int rows = left < kMeanRows ? static_cast<int>(left) : kMeanRows;
          const float* col = &ring[t & 1][lane][0];
          int r = 0;
          for (; r + 4 <= rows; r += 4) {
            const float4 q = *reinterpret_cast<const float4*>(col + r);
            m = m + q.x;
            m = m + q.y;
            m = m + q.z;
            m = m + q.w;
          }
          for (; r < rows; ++r) m = m + col[r];
        }
      }
      __syncthreads();
    }
  }
  if (wave == 0 && dd < d) mean[dd] = m / static_cast<float>(n);
}

// dm: dimensions that enter the max-abs and get scaled (the reference:
// min(D, 256), .cu:57-65, :76-84; option scale_all: D)
__global__ __launch_bounds__(256) void center_kernel(float* __restrict__ x, int64_t n, int d, int ld, int dm,
                                                     const float* __restrict__ mean, uint32_t* __restrict__ gmax) {
  __shared__ float red[256];
  float local = 0.0f;
  const int64_t total = n * ld;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < total; e += gridDim.x * 256ll) {
    const int dd = static_cast<int>(e % ld);
    if (dd < d) {
      const float v = x[e] - mean[dd];
      x[e] = v;
      if (dd < dm) {
        const float a = __builtin_fabsf(v);
        local = a > local ? a : local;
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

__global__ __launch_bounds__(256) void scale_kernel(float* __restrict__ x, int64_t n, int d, int ld, int dm,
                                                    const uint32_t* __restrict__ gmax) {
  const float g = __uint_as_float(*gmax);
  const int64_t total = n * ld;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < total; e += gridDim.x * 256ll) {
    const int dd = static_cast<int>(e % ld);
    if (dd < dm) x[e] = x[e] / g;
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

void launch_normalize(float* x, int64_t n, int d, int ld, float* mean, uint32_t* gmax, bool scale_all,
                      hipStream_t s) {
  const int dm = scale_all ? d : (d < 256 ? d : 256);
  AKNN_HIP_TRY(hipMemsetAsync(gmax, 0, sizeof(uint32_t), s));
  hipLaunchKernelGGL(mean_kernel, dim3((d + 63) / 64), dim3(1024), 0, s, x, n, d, ld, mean);
  AKNN_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(center_kernel, dim3(grid_for(n * ld, 256, 8192)), dim3(256), 0, s, x, n, d, ld, dm, mean, gmax);
  AKNN_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n * ld, 256, 8192)), dim3(256), 0, s, x, n, d, ld, dm, gmax);
  AKNN_HIP_TRY(hipGetLastError());
}

void launch_norms(const float* x, int64_t n, int d, int ld, float* norms, hipStream_t s) {
  hipLaunchKernelGGL(norms_kernel, dim3(grid_for(n, 256, 1 << 30)), dim3(256), 0, s, x, n, d, ld, norms);
  AKNN_HIP_TRY(hipGetLastError());
}

}  // namespace aknn
