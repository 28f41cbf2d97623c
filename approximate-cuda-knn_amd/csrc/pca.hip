// Principal-subspace screen of the propagation candidates.
//
// For P with orthonormal-ish rows, ||P(x - y)|| <= sigma ||x - y|| with sigma
// an upper bound of P's largest singular value, so the distance of the
// points' projections gives a proven lower bound of their distance.  Each
// point keeps its projection p(x) = fl(P x) (r = 16 or 32 FP32 values, one FMA
// chain per coordinate in dimension order) and a radius e(x) >= ||p(x) - P x||
// bounding the chains' rounding.  For data near an r-dimensional subspace
// (C5: a 16-dimensional mixture embedded in 128 dimensions) the bound is
// nearly the distance itself, and the screen reads r*4 bytes per candidate
// instead of its 4*D-byte row; survivors take the exact chain, so the graph is
// the same bit for bit.  P holds the top eigenvectors of a sample covariance
// (any P is valid; a better one is only tighter), recomputed on every build:
// the sample's second moments on the device, a subspace iteration in FP64 on
// the host.
#include <algorithm>
#include <cmath>
#include <vector>

#include "device_util.hpp"
#include "directions.hpp"
#include "kernels.hpp"

namespace aknn {

namespace {

constexpr int kMomentChunks = 32;

// second moments of a sample (rows q*stride, q < m) for pairs a <= b, FP32:
// blockIdx.y takes one of kMomentChunks consecutive row ranges and writes its
// partial sums; moments_reduce_kernel adds the chunks in order
__global__ __launch_bounds__(256) void moments_kernel(const float* __restrict__ x, int d, int ld, int64_t stride,
                                                      int64_t m, float* __restrict__ part) {
  const int64_t e = blockIdx.x * 256ll + threadIdx.x;
  if (e >= static_cast<int64_t>(d) * d) return;
  const int a = static_cast<int>(e / d), b = static_cast<int>(e % d);
  if (a > b) return;
  const int64_t per = (m + kMomentChunks - 1) / kMomentChunks;
  const int64_t q0 = blockIdx.y * per, q1 = q0 + per < m ? q0 + per : m;
  float s = 0.0f;
  for (int64_t q = q0; q < q1; ++q) {
    const float* r = x + q * stride * ld;
    s = __builtin_fmaf(r[a], r[b], s);
  }
  part[static_cast<int64_t>(blockIdx.y) * d * d + a * d + b] = s;
}

__global__ __launch_bounds__(256) void moments_reduce_kernel(const float* __restrict__ part, int d,
                                                             float* __restrict__ out) {
  const int64_t e = blockIdx.x * 256ll + threadIdx.x;
  if (e >= static_cast<int64_t>(d) * d) return;
  const int a = static_cast<int>(e / d), b = static_cast<int>(e % d);
  if (a > b) return;
  float s = 0.0f;
  for (int c = 0; c < kMomentChunks; ++c) s += part[static_cast<int64_t>(c) * d * d + a * d + b];
  out[a * d + b] = s;
  out[b * d + a] = s;
}

// p(x) and its radius: E_j = g * sum_d |P_jd x_d| bounds the rounding of
// coordinate j's chain (g = 2 (D + 8) 2^-24 covers the D-term FMA chain with a
// factor 2 to spare); the sum of magnitudes is itself an FP32 FMA chain of
// non-negative terms, at least (1 - g/2) of the exact one, so it is scaled by
// (1 + g); e = ||E|| is evaluated in FP64 and rounded up, plus the absolute
// subnormal term.
template <int R>
__global__ __launch_bounds__(256) void pca_project_kernel(const float* __restrict__ x, int64_t n, int d, int ld,
                                                          const float* __restrict__ p, float* __restrict__ proj,
                                                          _Float16* __restrict__ proj16, float* __restrict__ rad,
                                                          double g, double tiny) {
  __shared__ float ps[R * 256];
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  const float* xr = x + (i < n ? i : 0) * ld;
  float acc[R], ab[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    acc[j] = 0.0f;
    ab[j] = 0.0f;
  }
  for (int c0 = 0; c0 < d; c0 += 256) {
    const int cw = d - c0 < 256 ? d - c0 : 256;
    __syncthreads();
    for (int e = threadIdx.x; e < R * 256; e += 256) {
      const int j = e / 256, c = e % 256;
      ps[e] = c < cw ? p[static_cast<int64_t>(j) * d + c0 + c] : 0.0f;
    }
    __syncthreads();
    // 16-byte row loads, eight in flight per lane before their FMAs
    int c = 0;
    for (; c + 32 <= cw; c += 32) {
      float4 q[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) q[u] = *reinterpret_cast<const float4*>(xr + c0 + c + 4 * u);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const float f[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const float av = __builtin_fabsf(f[t]);
#pragma unroll
          for (int j = 0; j < R; ++j) {
            const float pv = ps[j * 256 + c + 4 * u + t];
            acc[j] = __builtin_fmaf(f[t], pv, acc[j]);
            ab[j] = __builtin_fmaf(av, __builtin_fabsf(pv), ab[j]);
          }
        }
      }
    }
    for (; c < cw; ++c) {
      const float v = xr[c0 + c];
      const float av = __builtin_fabsf(v);
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const float pv = ps[j * 256 + c];
        acc[j] = __builtin_fmaf(v, pv, acc[j]);
        ab[j] = __builtin_fmaf(av, __builtin_fabsf(pv), ab[j]);
      }
    }
  }
  if (i < n) {
    double e2 = 0.0, h2 = 0.0;
    bool over = false;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      if (proj16 != nullptr) {
        // FP16 rounding of the coordinate, its error kept exactly (FP64
        // holds the difference of an FP32 and an FP16 value)
        const _Float16 h = static_cast<_Float16>(acc[j]);
        const double hd = static_cast<double>(static_cast<float>(h)) - static_cast<double>(acc[j]);
        over |= !(__builtin_fabsf(static_cast<float>(h)) <= 65504.0f);
        h2 += hd * hd;
        proj16[i * R + j] = h;
      } else {
        proj[i * R + j] = acc[j];
      }
      const double ej = g * static_cast<double>(ab[j]) * (1.0 + g) + tiny;
      e2 += ej * ej;
    }
    // ||h - P x|| <= ||h - p|| + ||p - P x||
    const double e = over ? __builtin_inf() : (sqrt(e2) + sqrt(h2)) * (1.0 + 1e-12);
    float f = static_cast<float>(e);
    if (static_cast<double>(f) < e) f = __uint_as_float(__float_as_uint(f) + 1u);
    rad[i] = f;
  }
}

// Eigenpairs of a small symmetric matrix (cyclic Jacobi, FP64), descending.
void jacobi_eigen(std::vector<double>& a, int q, std::vector<double>& w, std::vector<double>& v) {
  v.assign(static_cast<size_t>(q) * q, 0.0);
  for (int i = 0; i < q; ++i) v[i * q + i] = 1.0;
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < q; ++i)
      for (int j = i + 1; j < q; ++j) off += a[i * q + j] * a[i * q + j];
    if (off < 1e-30) break;
    for (int pi = 0; pi < q; ++pi)
      for (int qi = pi + 1; qi < q; ++qi) {
        const double apq = a[pi * q + qi];
        if (std::fabs(apq) < 1e-300) continue;
        const double th = (a[qi * q + qi] - a[pi * q + pi]) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < q; ++k) {
          const double akp = a[k * q + pi], akq = a[k * q + qi];
          a[k * q + pi] = c * akp - s * akq;
          a[k * q + qi] = s * akp + c * akq;
        }
        for (int k = 0; k < q; ++k) {
          const double apk = a[pi * q + k], aqk = a[qi * q + k];
          a[pi * q + k] = c * apk - s * aqk;
          a[qi * q + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < q; ++k) {
          const double vkp = v[k * q + pi], vkq = v[k * q + qi];
          v[k * q + pi] = c * vkp - s * vkq;
          v[k * q + qi] = s * vkp + c * vkq;
        }
      }
  }
  w.resize(q);
  for (int i = 0; i < q; ++i) w[i] = a[i * q + i];
}

}  // namespace

bool launch_pca_moments(const float* x, int64_t n, int d, int ld, float* moments_dev, hipStream_t s) {
  if (d < 16 || d > 256 || n < 64) return false;
  const int64_t m = std::min<int64_t>(n, 16384);
  const int64_t stride = n / m;
  const unsigned pg = static_cast<unsigned>((static_cast<int64_t>(d) * d + 255) / 256);
  float* part = moments_dev + static_cast<int64_t>(d) * d;
  hipLaunchKernelGGL(moments_kernel, dim3(pg, kMomentChunks), dim3(256), 0, s, x, d, ld, stride, m, part);
  AKNN_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(moments_reduce_kernel, dim3(pg), dim3(256), 0, s, part, d, moments_dev);
  AKNN_HIP_TRY(hipGetLastError());
  return true;
}

PcaBasis pca_basis_from_moments(const float* moments, int d) {
  PcaBasis b;
  std::vector<double> c(moments, moments + static_cast<size_t>(d) * d);
  double trace = 0.0;
  for (int i = 0; i < d; ++i) trace += c[static_cast<size_t>(i) * d + i];
  if (!(trace > 0.0)) return b;
  // subspace iteration for the top q = 32 eigenvectors (deterministic start)
  const int q = std::min(32, d);
  std::vector<double> qm(static_cast<size_t>(d) * q), z(static_cast<size_t>(d) * q);
  SplitMix64 gen(0x9c0ffee1234ull);
  for (auto& v : qm) v = gen.open01() - 0.5;
  auto orthonormalize = [&](std::vector<double>& a) {
    for (int j = 0; j < q; ++j) {
      for (int pj = 0; pj < j; ++pj) {
        double dot = 0.0;
        for (int i = 0; i < d; ++i) dot += a[i * q + j] * a[i * q + pj];
        for (int i = 0; i < d; ++i) a[i * q + j] -= dot * a[i * q + pj];
      }
      double nrm = 0.0;
      for (int i = 0; i < d; ++i) nrm += a[i * q + j] * a[i * q + j];
      nrm = std::sqrt(nrm);
      for (int i = 0; i < d; ++i) a[i * q + j] = nrm > 0 ? a[i * q + j] / nrm : (i == j ? 1.0 : 0.0);
    }
  };
  orthonormalize(qm);
  for (int it = 0; it < 8; ++it) {
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < q; ++j) {
        double sacc = 0.0;
        for (int k = 0; k < d; ++k) sacc += c[static_cast<size_t>(i) * d + k] * qm[k * q + j];
        z[i * q + j] = sacc;
      }
    orthonormalize(z);
    qm.swap(z);
  }
  // Rayleigh-Ritz: T = Q^T C Q, its eigenvectors rotate Q
  std::vector<double> cq(static_cast<size_t>(d) * q), t(static_cast<size_t>(q) * q), w, vv;
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < q; ++j) {
      double sacc = 0.0;
      for (int k = 0; k < d; ++k) sacc += c[static_cast<size_t>(i) * d + k] * qm[k * q + j];
      cq[i * q + j] = sacc;
    }
  for (int a = 0; a < q; ++a)
    for (int bb = 0; bb < q; ++bb) {
      double sacc = 0.0;
      for (int i = 0; i < d; ++i) sacc += qm[i * q + a] * cq[i * q + bb];
      t[a * q + bb] = sacc;
    }
  jacobi_eigen(t, q, w, vv);
  std::vector<int> ord(q);
  for (int i = 0; i < q; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int a, int bb) { return w[a] > w[bb]; });
  // retained share of the sample's second moment: 16 or 32 directions
  double cum = 0.0;
  int r = 0;
  for (int j = 0; j < q; ++j) {
    cum += std::max(0.0, w[ord[j]]);
    if ((j + 1 == 16 || j + 1 == 32) && cum >= 0.97 * trace) {
      r = j + 1;
      break;
    }
  }
  b.retained16 = 0.0;
  for (int j = 0; j < std::min(16, q); ++j) b.retained16 += std::max(0.0, w[ord[j]]) / trace;
  if (r == 0) return b;
  b.r = r;
  b.p.assign(static_cast<size_t>(r) * d, 0.0f);
  for (int j = 0; j < r; ++j)
    for (int i = 0; i < d; ++i) {
      double sacc = 0.0;
      for (int k = 0; k < q; ++k) sacc += qm[i * q + k] * vv[k * q + ord[j]];
      b.p[static_cast<size_t>(j) * d + i] = static_cast<float>(sacc);
    }
  // sigma^2 <= max row sum of |P P^T| (Gershgorin), from the FP32 entries in FP64
  double gmax = 0.0;
  for (int a = 0; a < r; ++a) {
    double row = 0.0;
    for (int bb = 0; bb < r; ++bb) {
      double sacc = 0.0;
      for (int i = 0; i < d; ++i)
        sacc += static_cast<double>(b.p[static_cast<size_t>(a) * d + i]) * b.p[static_cast<size_t>(bb) * d + i];
      row += std::fabs(sacc);
    }
    gmax = std::max(gmax, row);
  }
  b.sigma = std::sqrt(gmax * (1.0 + 1e-9)) * (1.0 + 1e-12);
  return b;
}

void launch_pca_project(const float* x, int64_t n, int d, int ld, const float* p, int r, float* proj, float* rad,
                        hipStream_t s, _Float16* proj16) {
  const double g = 2.0 * (d + 8) / 16777216.0;
  const double tiny = (d + 16) * 1.401298464324817e-45;
  const unsigned grid = static_cast<unsigned>((n + 255) / 256);
  if (r == 16)
    hipLaunchKernelGGL(pca_project_kernel<16>, dim3(grid), dim3(256), 0, s, x, n, d, ld, p, proj, proj16, rad, g,
                       tiny);
  else if (r == 32)
    hipLaunchKernelGGL(pca_project_kernel<32>, dim3(grid), dim3(256), 0, s, x, n, d, ld, p, proj, proj16, rad, g,
                       tiny);
  else
    throw HipError("pca: r must be 16 or 32");
  AKNN_HIP_TRY(hipGetLastError());
}

}  // namespace aknn
