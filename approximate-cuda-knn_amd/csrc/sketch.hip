// Half-precision sketch of the normalized points, with a per-point error radius.
//
// xh[i] is x[i] rounded to FP16 (round to nearest even) with values below the
// smallest normal FP16 (2^-14) flushed to zero, so no FP16 subnormal is ever
// stored; rows are padded with zeros to ldh (a multiple of 8 halves, 16 B).
// eps[i] is an upper bound of ||x[i] - xh[i]|| in the norm the distance needs
// (L1 for TAXICAB, L2 otherwise): every difference x - xh is exact in FP32,
// squares and sums are taken in FP64 and the result is rounded up to FP32.
// nh[i] is ||xh[i]||^2 rounded to nearest FP32 from its exact FP64 sum (the
// squares of halves are exact in FP64).  A row that does not fit FP16
// (|x| >= 65520, or NaN) gets eps = +inf: every bound built from it lets the
// exact path decide.
//
// The sketch feeds lower bounds of the exact FP32 keys (device_util.hpp,
// sketch_lower_bound): a candidate whose bound is above the current k-th key
// cannot enter, so skipping its exact distance changes no result.
#include "device_util.hpp"
#include "kernels.hpp"

namespace aknn {

namespace {

constexpr int kSketchWaves = 4;

__global__ __launch_bounds__(256) void sketch_kernel(const float* __restrict__ x, int64_t n, int d, int ld,
                                                     int l1, _Float16* __restrict__ xh, int ldh,
                                                     float* __restrict__ eps, float* __restrict__ nh) {
  const int lane = threadIdx.x & 63;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kSketchWaves + (threadIdx.x >> 6);
  if (i >= n) return;
  const float* xr = x + i * ld;
  _Float16* hr = xh + i * ldh;
  double acc = 0.0, sq = 0.0;
  bool bad = false;
  for (int dd = lane; dd < ldh; dd += 64) {
    const float v = dd < d ? xr[dd] : 0.0f;
    _Float16 h = static_cast<_Float16>(v);
    float hv = static_cast<float>(h);
    if (!(__builtin_fabsf(hv) <= 65504.0f)) {  // overflow to inf, or NaN input
      bad = true;
      h = static_cast<_Float16>(0.0f);
      hv = 0.0f;
    } else if (__builtin_fabsf(hv) < 6.103515625e-05f) {  // FP16 subnormal range: store zero
      h = static_cast<_Float16>(0.0f);
      hv = 0.0f;
    }
    hr[dd] = h;
    const double e = static_cast<double>(v) - static_cast<double>(hv);  // exact
    acc += l1 ? __builtin_fabs(e) : e * e;
    sq += static_cast<double>(hv) * static_cast<double>(hv);
  }
  for (int off = 32; off > 0; off >>= 1) {
    acc += __shfl_xor(acc, off);
    sq += __shfl_xor(sq, off);
  }
  bad = __any(bad);
  if (lane == 0) {
    // acc: D exact terms summed in FP64 (relative error below D * 2^-53);
    // widen by 1e-12, take the root for L2, round up to FP32
    double r = acc * (1.0 + 1e-12);
    if (!l1) r = __builtin_sqrt(r) * (1.0 + 1e-15);
    float f = static_cast<float>(r);
    if (static_cast<double>(f) < r) f = __uint_as_float(__float_as_uint(f) + 1u);  // r >= 0: next float up
    eps[i] = bad || !(f == f) ? __builtin_inff() : f;
    nh[i] = static_cast<float>(sq);
  }
}

}  // namespace

void launch_sketch(const float* x, int64_t n, int d, int ld, int metric, _Float16* xh, int ldh, float* eps,
                   float* nh, hipStream_t s) {
  if (ldh % 8 != 0 || ldh < d) throw HipError("sketch: row stride must be a multiple of 8 covering d");
  const int64_t blocks = (n + kSketchWaves - 1) / kSketchWaves;
  hipLaunchKernelGGL(sketch_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, x, n, d, ld,
                     metric == kTaxicab ? 1 : 0, xh, ldh, eps, nh);
  AKNN_HIP_TRY(hipGetLastError());
}

}  // namespace aknn
