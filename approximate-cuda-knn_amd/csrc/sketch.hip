// Sketch side products: the int8 sketch (propagation's integer screen) and
// the largest radius and norm of a sketch.  The FP16 sketch itself
// (device_util.hpp, sketch_element) is formed by normalization's row pass.
//
// The sketches feed lower bounds of the exact FP32 keys (device_util.hpp,
// sketch_lower_bound): a candidate whose bound is above the current k-th key
// cannot enter, so skipping its exact distance changes no result.
#include "device_util.hpp"
#include "kernels.hpp"

namespace aknn {

namespace {

constexpr int kSketchWaves = 4;

// Int8 sketch: q[i] = round(x[i] / s_i) with s_i = max_d |x_i,d| / 127, so
// x[i] ~ s_i q[i]; a8[i] = sum q^2 (exact), e8[i] an upper bound of
// ||x[i] - s_i q[i]||_2 (FP64, rounded up).  Rows are padded with zeros to
// ld8 bytes (a multiple of 16).  Dot products of two such rows are exact
// integers (v_dot4_i32_i8), so a distance bound needs only the two radii.
__global__ __launch_bounds__(256) void sketch8_kernel(const float* __restrict__ x, int64_t n, int d, int ld,
                                                      int8_t* __restrict__ q8, int ld8, float* __restrict__ s8,
                                                      int32_t* __restrict__ a8, float* __restrict__ e8,
                                                      float2* __restrict__ se) {
  const int lane = threadIdx.x & 63;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kSketchWaves + (threadIdx.x >> 6);
  if (i >= n) return;
  const float* xr = x + i * ld;
  float mx = 0.0f;
  bool bad = false;
  for (int dd = lane; dd < d; dd += 64) {
    const float v = xr[dd];
    if (!(__builtin_fabsf(v) <= 3.0e38f)) bad = true;
    mx = __builtin_fmaxf(mx, __builtin_fabsf(v));
  }
  for (int off = 32; off > 0; off >>= 1) mx = __builtin_fmaxf(mx, __shfl_xor(mx, off));
  bad = __any(bad);
  const float sc = bad || mx == 0.0f ? 0.0f : mx / 127.0f;
  const float inv = sc > 0.0f ? 127.0f / mx : 0.0f;
  double err = 0.0;
  int sq = 0;
  int8_t* qr = q8 + i * ld8;
  for (int dd = lane; dd < ld8; dd += 64) {
    int q = 0;
    double e = 0.0;
    if (dd < d) {
      const float v = bad ? 0.0f : xr[dd];
      q = static_cast<int>(__builtin_rintf(v * inv));
      q = q > 127 ? 127 : (q < -127 ? -127 : q);
      e = static_cast<double>(v) - static_cast<double>(q) * static_cast<double>(sc);  // exact
    }
    qr[dd] = static_cast<int8_t>(q);
    err += e * e;
    sq += q * q;
  }
  for (int off = 32; off > 0; off >>= 1) {
    err += __shfl_xor(err, off);
    sq += __shfl_xor(sq, off);
  }
  if (lane == 0) {
    const double r = __builtin_sqrt(err * (1.0 + 1e-12)) * (1.0 + 1e-15);
    float f = static_cast<float>(r);
    if (static_cast<double>(f) < r) f = __uint_as_float(__float_as_uint(f) + 1u);
    s8[i] = sc;
    a8[i] = sq;
    e8[i] = bad ? __builtin_inff() : f;
    if (se != nullptr) se[i] = make_float2(sc, bad ? __builtin_inff() : f);
  }
}

}  // namespace

void launch_sketch8(const float* x, int64_t n, int d, int ld, int8_t* q8, int ld8, float* s8, int32_t* a8, float* e8,
                    float2* se, hipStream_t s) {
  if (ld8 % 16 != 0 || ld8 < d) throw HipError("sketch8: row stride must be a multiple of 16 covering d");
  const int64_t blocks = (n + kSketchWaves - 1) / kSketchWaves;
  hipLaunchKernelGGL(sketch8_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, x, n, d, ld, q8, ld8, s8,
                     a8, e8, se);
  AKNN_HIP_TRY(hipGetLastError());
}

namespace {

// out = max(v[0..n)) for non-negative v (+inf allowed): grid-stride partial
// maxima, one atomic per block (non-negative floats order like their bits).
__global__ __launch_bounds__(256) void max_nonneg_kernel(const float* __restrict__ v, int64_t n,
                                                         uint32_t* __restrict__ out) {
  __shared__ uint32_t red[256];
  uint32_t m = 0u;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += gridDim.x * 256ll) {
    const uint32_t b = __float_as_uint(v[e]);
    m = b > m ? b : m;
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s && red[threadIdx.x + s] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, red[0]);
}

}  // namespace

void launch_max_nonneg(const float* v, int64_t n, float* out, hipStream_t s) {
  AKNN_HIP_TRY(hipMemsetAsync(out, 0, sizeof(float), s));
  const int64_t g = (n + 255) / 256;
  hipLaunchKernelGGL(max_nonneg_kernel, dim3(static_cast<unsigned>(g < 256 ? (g < 1 ? 1 : g) : 256)), dim3(256), 0, s,
                     v, n, reinterpret_cast<uint32_t*>(out));
  AKNN_HIP_TRY(hipGetLastError());
}

}  // namespace aknn
