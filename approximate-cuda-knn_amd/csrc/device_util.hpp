// Device helpers shared by the kernels: the exact FP32 distance chains and
// wave64 utilities.
//
// Bit-exactness rules (the reference is nvcc-compiled with default FMA
// contraction and IEEE division/sqrt, euclideandistancematrixgpu.cu:243-273,
// :540-570, :840-881):
//   * every multiply-add of a chain is an explicit fmaf, every other operation
//     an explicit single operation, accumulated in dimension order;
//   * dimensions are processed in chunks of 256, each chunk's partial starting
//     at 0.0f and being added into the total;
//   * sqrtf and '/' are the correctly rounded forms (built with
//     -fhip-fp32-correctly-rounded-divide-sqrt), denormals are preserved.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "aknn_common.hpp"

namespace aknn {

constexpr int kWave = 64;

// (key, id) lexicographic order used by every selection: ties go to the lower id.
__device__ __forceinline__ bool key_less(float ka, int32_t ia, float kb, int32_t ib) {
  return ka < kb || (ka == kb && ia < ib);
}

// Final key of a pair from its accumulated totals.
template <int METRIC>
__device__ __forceinline__ float finish_key(float tot, float aa, float bb) {
  if constexpr (METRIC == kCosine) {
    const float c = tot / (sqrtf(aa) * sqrtf(bb));
    return static_cast<float>(1.0 - static_cast<double>(c));
  } else if constexpr (METRIC == kTaxicabL1) {
    return tot;
  } else {
    return sqrtf(tot);
  }
}

// One step of a chain: EUCLIDEAN fma(t, t, p), TAXICAB p + |t|, COSINE fma(a, b, p).
template <int METRIC>
__device__ __forceinline__ float chain_step(float a, float b, float p) {
  if constexpr (METRIC == kEuclidean) {
    const float t = a - b;
    return __builtin_fmaf(t, t, p);
  } else if constexpr (METRIC == kTaxicab || METRIC == kTaxicabL1) {
    return p + __builtin_fabsf(a - b);
  } else {
    return __builtin_fmaf(a, b, p);
  }
}

// Chain over dimensions [0, d) of row a (any address space) and row b
// (global, 16-byte aligned, padded to a multiple of 4).  Returns the total.
template <int METRIC, int UNROLL = 8>
__device__ __forceinline__ float row_chain(const float* a, const float* __restrict__ b, int d) {
  float tot = 0.0f;
  for (int c0 = 0; c0 < d; c0 += 256) {
    const int c1 = c0 + 256 < d ? c0 + 256 : d;
    float p = 0.0f;
    int dd = c0;
    // UNROLL 16-byte loads in flight per lane before their FMAs
#pragma unroll UNROLL
    for (; dd + 4 <= c1; dd += 4) {
      const float4 bv = *reinterpret_cast<const float4*>(b + dd);
      p = chain_step<METRIC>(a[dd + 0], bv.x, p);
      p = chain_step<METRIC>(a[dd + 1], bv.y, p);
      p = chain_step<METRIC>(a[dd + 2], bv.z, p);
      p = chain_step<METRIC>(a[dd + 3], bv.w, p);
    }
    for (; dd < c1; ++dd) p = chain_step<METRIC>(a[dd], b[dd], p);
    tot = tot + p;
  }
  return tot;
}

// The same chain with the row of b read U 16-byte loads at a time into
// registers, all issued before the first step that uses them (a scheduling
// barrier keeps the compiler from interleaving them with the steps, which
// it otherwise does to save registers): one memory round trip per 4U
// dimensions.  For kernels whose occupancy is bound by LDS, not registers.
template <int METRIC, int U>
__device__ __forceinline__ float row_chain_batched(const float* a, const float* __restrict__ b, int d) {
  float tot = 0.0f;
  for (int c0 = 0; c0 < d; c0 += 256) {
    const int c1 = c0 + 256 < d ? c0 + 256 : d;
    float p = 0.0f;
    int dd = c0;
    for (; dd + 4 * U <= c1; dd += 4 * U) {
      float4 v[U];
#pragma unroll
      for (int j = 0; j < U; ++j) v[j] = *reinterpret_cast<const float4*>(b + dd + 4 * j);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < U; ++j) {
        p = chain_step<METRIC>(a[dd + 4 * j + 0], v[j].x, p);
        p = chain_step<METRIC>(a[dd + 4 * j + 1], v[j].y, p);
        p = chain_step<METRIC>(a[dd + 4 * j + 2], v[j].z, p);
        p = chain_step<METRIC>(a[dd + 4 * j + 3], v[j].w, p);
      }
    }
    for (; dd + 4 <= c1; dd += 4) {
      const float4 bv = *reinterpret_cast<const float4*>(b + dd);
      p = chain_step<METRIC>(a[dd + 0], bv.x, p);
      p = chain_step<METRIC>(a[dd + 1], bv.y, p);
      p = chain_step<METRIC>(a[dd + 2], bv.z, p);
      p = chain_step<METRIC>(a[dd + 3], bv.w, p);
    }
    for (; dd < c1; ++dd) p = chain_step<METRIC>(a[dd], b[dd], p);
    tot = tot + p;
  }
  return tot;
}

// Same chain with both rows in global memory (16-byte aligned, padded to a
// multiple of 4): 16-byte loads from both, 2 * UNROLL in flight per lane.
template <int METRIC, int UNROLL = 4>
__device__ __forceinline__ float row_chain_global(const float* __restrict__ a, const float* __restrict__ b, int d) {
  float tot = 0.0f;
  for (int c0 = 0; c0 < d; c0 += 256) {
    const int c1 = c0 + 256 < d ? c0 + 256 : d;
    float p = 0.0f;
    int dd = c0;
#pragma unroll UNROLL
    for (; dd + 4 <= c1; dd += 4) {
      const float4 av = *reinterpret_cast<const float4*>(a + dd);
      const float4 bv = *reinterpret_cast<const float4*>(b + dd);
      p = chain_step<METRIC>(av.x, bv.x, p);
      p = chain_step<METRIC>(av.y, bv.y, p);
      p = chain_step<METRIC>(av.z, bv.z, p);
      p = chain_step<METRIC>(av.w, bv.w, p);
    }
    for (; dd < c1; ++dd) p = chain_step<METRIC>(a[dd], b[dd], p);
    tot = tot + p;
  }
  return tot;
}

// Same chain for exactly d = 4*NV <= 256 dimensions (one chunk), with the
// whole row of b loaded into registers before the first step: one memory
// latency per row instead of one per batch of loads.  Kernels are
// instantiated per row width (no in-kernel switch over widths), so each
// instantiation has a single straight-line load path.
template <int METRIC, int NV>
__device__ __forceinline__ float row_chain_regs(const float* a, const float* __restrict__ b) {
  float4 v[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] = *reinterpret_cast<const float4*>(b + 4 * j);
  float p = 0.0f;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    p = chain_step<METRIC>(a[4 * j + 0], v[j].x, p);
    p = chain_step<METRIC>(a[4 * j + 1], v[j].y, p);
    p = chain_step<METRIC>(a[4 * j + 2], v[j].z, p);
    p = chain_step<METRIC>(a[4 * j + 3], v[j].w, p);
  }
  return 0.0f + p;
}

// The FP16 sketch of a normalized point: xh[i] is x[i] rounded to FP16
// (round to nearest even) with values below the smallest normal FP16 (2^-14)
// flushed to zero, so no FP16 subnormal is ever stored; rows are padded with
// zeros to ldh (a multiple of 8 halves, 16 B).  eps[i] is an upper bound of
// ||x[i] - xh[i]|| in the norm the distance needs (L1 for TAXICAB, L2
// otherwise): every difference x - xh is exact in FP32, squares and sums are
// taken in FP64 and the result is rounded up to FP32.  nh[i] is ||xh[i]||^2
// rounded to nearest FP32 from its exact FP64 sum (the squares of halves are
// exact in FP64).  A row that does not fit FP16 (|x| >= 65520, or NaN) gets
// eps = +inf: every bound built from it lets the exact path decide.
//
// sketch_element: one element's half, its error term added to acc and its
// square to sq; bad set when the element does not fit.
__device__ __forceinline__ _Float16 sketch_element(float v, bool l1, double& acc, double& sq, bool& bad) {
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
  const double e = static_cast<double>(v) - static_cast<double>(hv);  // exact
  acc += l1 ? __builtin_fabs(e) : e * e;
  sq += static_cast<double>(hv) * static_cast<double>(hv);
  return h;
}

// The radius from a row's summed error terms: D exact terms summed in FP64
// (relative error below D * 2^-53), widened by 1e-12, the root taken for L2,
// rounded up to FP32; +inf for a row that does not fit.
__device__ __forceinline__ float sketch_radius(double acc, bool l1, bool bad) {
  double r = acc * (1.0 + 1e-12);
  if (!l1) r = __builtin_sqrt(r) * (1.0 + 1e-15);
  float f = static_cast<float>(r);
  if (static_cast<double>(f) < r) f = __uint_as_float(__float_as_uint(f) + 1u);  // r >= 0: next float up
  return bad || !(f == f) ? __builtin_inff() : f;
}

// Chain of a FP16 sketch row b (global, ldh halves, 16-byte aligned, zero
// padded) against a query sketch a held as floats (zero padded to ldh): one
// unchunked chain in FP32, each half converted exactly.  Feeds
// sketch_lower_bound only, never a stored value.
__device__ __forceinline__ float half_lo(uint32_t w) {
  return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w & 0xffffu)));
}
__device__ __forceinline__ float half_hi(uint32_t w) {
  return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w >> 16)));
}

// G 16-byte loads (8 halves each) issued together, then their 8G steps.
template <int METRIC, int G>
__device__ __forceinline__ float sketch_group(const float* a, const uint4* __restrict__ b, float p) {
  uint4 v[G];
#pragma unroll
  for (int j = 0; j < G; ++j) v[j] = b[j];
#pragma unroll
  for (int j = 0; j < G; ++j) {
    const float* aj = a + 8 * j;
    p = chain_step<METRIC>(aj[0], half_lo(v[j].x), p);
    p = chain_step<METRIC>(aj[1], half_hi(v[j].x), p);
    p = chain_step<METRIC>(aj[2], half_lo(v[j].y), p);
    p = chain_step<METRIC>(aj[3], half_hi(v[j].y), p);
    p = chain_step<METRIC>(aj[4], half_lo(v[j].z), p);
    p = chain_step<METRIC>(aj[5], half_hi(v[j].z), p);
    p = chain_step<METRIC>(aj[6], half_lo(v[j].w), p);
    p = chain_step<METRIC>(aj[7], half_hi(v[j].w), p);
  }
  return p;
}

template <int METRIC>
__device__ __forceinline__ float sketch_chain(const float* a, const _Float16* __restrict__ b, int ldh) {
  const uint4* bv = reinterpret_cast<const uint4*>(b);
  float p = 0.0f;
  int dd = 0;
  // 128 bytes in flight per lane (a whole row in flight measured slower: the
  // extra registers cost more waves than the deeper prefetch gains)
  for (; dd + 64 <= ldh; dd += 64) p = sketch_group<METRIC, 8>(a + dd, bv + dd / 8, p);
  for (; dd < ldh; dd += 8) p = sketch_group<METRIC, 1>(a + dd, bv + dd / 8, p);
  return p;
}

// Constants of the sketch bounds for rows of d dimensions.  kap bounds the
// relative error of any FP32 chain of the pipeline over d terms (the exact
// chunked chains take at most d + d/256 + 3 roundings per term, the sketch
// chain d + 1) with a factor 2 to spare, which also covers the FP64 rounding
// of the bound itself; tiny bounds the absolute error of subnormal results.
struct SketchBound {
  double kap;
  double tiny;
};
inline SketchBound sketch_bound(int d) {
  const double u = 1.0 / 16777216.0;  // 2^-24
  return {2.0 * (d + d / 256 + 8) * u, (d + 16) * 1.401298464324817e-45};
}

// A lower bound of the exact FP32 key (finish_key of the exact chain) of the
// pair (x, y), from the sketch chain p of (xh, yh), the sketch radii ex, ey
// (L2; L1 for TAXICAB) and, for COSINE, the exact squared-norm chains aa, bb.
//   EUCLIDEAN: ||x-y|| >= sqrt(p (1-kap)) - ex - ey (triangle inequality), and
//     the exact key >= ||x-y|| (1-kap).
//   TAXICAB: |x-y|_1 >= p (1-kap) - ex - ey, and key >= sqrt(|x-y|_1 (1-kap));
//     without the root (kTaxicabL1) key >= |x-y|_1 (1-kap).
//   COSINE: the exact key is 1 - fl(tot / den) with den the same FP32 product
//     of roots as finish_key, and tot <= p + kap (A+ex)(B+ey) + ex B
//     + (A+ex) ey + kap A B, with A >= ||x||, B >= ||y|| from aa, bb.
// Non-finite or degenerate inputs give -inf: the exact path then decides.
template <int METRIC>
__device__ __forceinline__ double sketch_lower_bound(float p, float ex, float ey, float aa, float bb,
                                                     const SketchBound& sb) {
  const double k1 = 1.0 - sb.kap;
  double lb;
  if constexpr (METRIC == kEuclidean) {
    const double s = fmax(0.0, static_cast<double>(p) * k1 - sb.tiny);
    const double r = fmax(0.0, sqrt(s) - static_cast<double>(ex) - static_cast<double>(ey));
    lb = sqrt(fmax(0.0, r * r * k1 - sb.tiny)) * k1;
  } else if constexpr (METRIC == kTaxicab) {
    const double s =
        fmax(0.0, static_cast<double>(p) * k1 - sb.tiny - static_cast<double>(ex) - static_cast<double>(ey));
    lb = sqrt(fmax(0.0, s * k1 - sb.tiny)) * k1;
  } else if constexpr (METRIC == kTaxicabL1) {
    // the exact L1 chain is at least |x-y|_1 (1-kap) - tiny
    const double s =
        fmax(0.0, static_cast<double>(p) * k1 - sb.tiny - static_cast<double>(ex) - static_cast<double>(ey));
    lb = fmax(0.0, s * k1 - sb.tiny);
  } else {
    const float den = sqrtf(aa) * sqrtf(bb);
    if (!(den > 0.0f) || !(den <= 3.4e38f)) return -__builtin_inf();
    const double a = sqrt((static_cast<double>(aa) + sb.tiny) * (1.0 + 2.0 * sb.kap));
    const double b = sqrt((static_cast<double>(bb) + sb.tiny) * (1.0 + 2.0 * sb.kap));
    const double xe = ex, ye = ey;
    const double up = static_cast<double>(p) + sb.kap * (a + xe) * (b + ye) + xe * b + (a + xe) * ye +
                      sb.kap * a * b + 2.0 * sb.tiny;
    double q = up / static_cast<double>(den);
    q += fabs(q) * 2.384185791015625e-07 + 1e-40;  // the FP32 quotient's rounding, 2^-22 relative
    lb = (1.0 - q) - 9.5367431640625e-07;          // the FP32 store of 1 - quotient, 2^-20 absolute
  }
  return lb == lb ? lb : -__builtin_inf();
}

// An upper bound of the exact COSINE key 1 - fl(tot / den) from the same
// terms: tot >= p - (kap (A+ex)(B+ey) + ex B + (A+ex) ey + kap A B), the
// error of the sketch's dot product being bounded on both sides alike.
// Non-finite or degenerate inputs give +inf.
__device__ __forceinline__ double sketch_upper_bound_cosine(float p, float ex, float ey, float aa, float bb,
                                                            const SketchBound& sb) {
  const float den = sqrtf(aa) * sqrtf(bb);
  if (!(den > 0.0f) || !(den <= 3.4e38f)) return __builtin_inf();
  const double a = sqrt((static_cast<double>(aa) + sb.tiny) * (1.0 + 2.0 * sb.kap));
  const double b = sqrt((static_cast<double>(bb) + sb.tiny) * (1.0 + 2.0 * sb.kap));
  const double xe = ex, ye = ey;
  const double lo = static_cast<double>(p) - sb.kap * (a + xe) * (b + ye) - xe * b - (a + xe) * ye -
                    sb.kap * a * b - 2.0 * sb.tiny;
  double q = lo / static_cast<double>(den);
  q -= fabs(q) * 2.384185791015625e-07 + 1e-40;  // the FP32 quotient's rounding, 2^-22 relative
  const double ub = (1.0 - q) + 9.5367431640625e-07;  // the FP32 store of 1 - quotient, 2^-20 absolute
  return ub == ub ? ub : __builtin_inf();
}

// The float next to f below (above); +inf steps to the largest finite float.
__device__ __forceinline__ float float_below(float f) {
  const uint32_t b = __float_as_uint(f);
  if (f > 0.0f) return __uint_as_float(b - 1u);
  if (f == 0.0f) return -1.401298464e-45f;
  return __uint_as_float(b + 1u);
}
__device__ __forceinline__ float float_above(float f) { return -float_below(-f); }

// The largest float not above v (the smallest not below v).
__device__ __forceinline__ float float_down(double v) {
  const float f = static_cast<float>(v);
  return static_cast<double>(f) > v ? float_below(f) : f;
}
__device__ __forceinline__ float float_up(double v) {
  const float f = static_cast<float>(v);
  return static_cast<double>(f) < v ? float_above(f) : f;
}

// Row-width class of a kernel instantiation: rows of exactly 4*NV floats in
// registers (NV = 8, 16, 24, 32), or 0 for any width, streamed.
inline int row_width_class(int d) {
  switch (d) {
    case 32: return 8;
    case 64: return 16;
    case 96: return 24;
    case 128: return 32;
    default: return 0;
  }
}

// Per-point squared norm with the reference's COSINE chain (fma(a, a, p)
// chunked by 256); the reference recomputes it for every pair, the value is
// the same every time.
__device__ __forceinline__ float norm_chain(const float* __restrict__ a, int d) {
  float tot = 0.0f;
  for (int c0 = 0; c0 < d; c0 += 256) {
    const int c1 = c0 + 256 < d ? c0 + 256 : d;
    float p = 0.0f;
    for (int dd = c0; dd < c1; ++dd) p = __builtin_fmaf(a[dd], a[dd], p);
    tot = tot + p;
  }
  return tot;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2:
// block b runs on XCD b % 8.  This maps XCD x to the contiguous logical range
// [x * grid/8, (x+1) * grid/8), so neighbouring work (the owner chunks of one
// leaf, spatially close points) shares one L2.  grid must be a multiple of 8.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t grid) {
  return (b & 7) * (grid >> 3) + (b >> 3);
}
inline int64_t xcd_grid(int64_t blocks) { return (blocks + 7) / 8 * 8; }

// Inserts (key, id) into the sorted list held by lanes [0, k) of the wave
// (lane l holds entry l) and drops the last entry.  The caller guarantees the
// new entry is strictly below the current last entry and not yet present.
__device__ __forceinline__ void wave_list_insert(float& key, int32_t& id, int k, float nk, int32_t nid) {
  const int lane = lane_id();
  const bool smaller = lane < k && key_less(key, id, nk, nid);
  const int pos = __popcll(__ballot(smaller));
  const float pk = __shfl_up(key, 1);
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

// Bitonic sort of the (key, id) pairs held one per lane over the whole wave
// (ascending; lanes beyond the list must hold +inf / INT32_MAX).
__device__ __forceinline__ void wave_sort(float& key, int32_t& id) {
  const int lane = lane_id();
  for (int size = 2; size <= kWave; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const float ok = __shfl_xor(key, stride);
      const int32_t oid = __shfl_xor(id, stride);
      const bool up = (lane & size) == 0;
      const bool lower = (lane & stride) == 0;
      const bool other_less = key_less(ok, oid, key, id);
      // lower lane keeps the min when sorting up, the max when sorting down
      const bool take = lower == up ? other_less : !other_less;
      if (take) {
        key = ok;
        id = oid;
      }
    }
  }
}

}  // namespace aknn
