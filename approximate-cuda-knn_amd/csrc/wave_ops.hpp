// Wave-level building blocks shared by the selection kernels: DPP scans
// and reductions over the 64 lanes, lane broadcasts, a bitonic sort of one
// 64-bit value per lane, and a radix select of the first k of a list of
// packed (key, id) pairs by one wave.
#pragma once

#include <cstdint>

namespace aknn {

// Inclusive prefix sum over the 64 lanes by DPP row shifts and row
// broadcasts (no LDS): within each row of 16 lanes by shifts of 1, 2, 4 and
// 8, then row 0's total into row 1 and row 2's into row 3 (row_bcast:15),
// then lane 31's into rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ int wave_incl_add(int v) {
  const int lane = threadIdx.x & 63, rl = lane & 15;
  int t;
  t = __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  if (rl >= 1) v += t;
  t = __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  if (rl >= 2) v += t;
  t = __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  if (rl >= 4) v += t;
  t = __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  if (rl >= 8) v += t;
  t = __builtin_amdgcn_update_dpp(0, v, 0x142, 0xf, 0xf, false);  // row_bcast:15
  if ((lane & 31) >= 16) v += t;
  t = __builtin_amdgcn_update_dpp(0, v, 0x143, 0xf, 0xf, false);  // row_bcast:31
  if (lane >= 32) v += t;
  return v;
}

// AND and OR of one 32-bit value over the 64 lanes (the same DPP steps,
// the result read from lane 63).
__device__ __forceinline__ uint32_t wave_and(uint32_t v) {
  const int lane = threadIdx.x & 63, rl = lane & 15;
  uint32_t t;
  t = __builtin_amdgcn_update_dpp(~0u, v, 0x111, 0xf, 0xf, false);
  if (rl >= 1) v &= t;
  t = __builtin_amdgcn_update_dpp(~0u, v, 0x112, 0xf, 0xf, false);
  if (rl >= 2) v &= t;
  t = __builtin_amdgcn_update_dpp(~0u, v, 0x114, 0xf, 0xf, false);
  if (rl >= 4) v &= t;
  t = __builtin_amdgcn_update_dpp(~0u, v, 0x118, 0xf, 0xf, false);
  if (rl >= 8) v &= t;
  t = __builtin_amdgcn_update_dpp(~0u, v, 0x142, 0xf, 0xf, false);
  if ((lane & 31) >= 16) v &= t;
  t = __builtin_amdgcn_update_dpp(~0u, v, 0x143, 0xf, 0xf, false);
  if (lane >= 32) v &= t;
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}
__device__ __forceinline__ uint32_t wave_or(uint32_t v) { return ~wave_and(~v); }

// Orders the LDS accesses of the lanes of one wave.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Lane j's 64-bit value, broadcast (j uniform); the halves are read as
// unsigned 32-bit values so that the low half does not sign-extend.
__device__ __forceinline__ uint64_t lane_u64(uint64_t v, int j) {
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), j));
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v)), j));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Bitonic sort of one 64-bit value per lane over the wave, ascending.
__device__ __forceinline__ uint64_t wave_sort_u64(uint64_t v) {
  const int lane = threadIdx.x & 63;
  for (int size = 2; size <= 64; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const uint32_t ohi = __shfl_xor(static_cast<uint32_t>(v >> 32), stride);
      const uint32_t olo = __shfl_xor(static_cast<uint32_t>(v), stride);
      const uint64_t o = (static_cast<uint64_t>(ohi) << 32) | olo;
      const bool up = (lane & size) == 0, lower = (lane & stride) == 0;
      const bool take = lower == up ? o < v : !(o < v);
      v = take ? o : v;
    }
  }
  return v;
}

// Count of the entries of the ascending array a[0, len) (len <= 64, LDS)
// below v.
__device__ __forceinline__ int count_below(const uint64_t* a, int len, uint64_t v) {
  int c = 0;
#pragma unroll
  for (int step = 32; step >= 1; step >>= 1)
    if (c + step <= len && a[c + step - 1] < v) c += step;
  return c;
}

// First min(k, n) of the distinct packed pairs src[0, n) (LDS) by one wave,
// in any order, into out: a radix select from the highest bit on which the
// entries differ (the bits above it are common to all), up to 8 bits per
// round, over a per-wave histogram wh[256].  Returns the count taken.
__device__ inline int wave_topk(const uint64_t* src, int n, int k, uint64_t* out, uint32_t* wh) {
  const int lane = threadIdx.x & 63;
  if (n <= k) {
    for (int e = lane; e < n; e += 64) out[e] = src[e];
    return n;
  }
  uint32_t alo = ~0u, ahi = ~0u, olo = 0u, ohi = 0u;
  for (int e = lane; e < n; e += 64) {
    const uint64_t v = src[e];
    alo &= static_cast<uint32_t>(v);
    ahi &= static_cast<uint32_t>(v >> 32);
    olo |= static_cast<uint32_t>(v);
    ohi |= static_cast<uint32_t>(v >> 32);
  }
  alo = wave_and(alo);
  ahi = wave_and(ahi);
  olo = wave_or(olo);
  ohi = wave_or(ohi);
  const uint64_t all_and = (static_cast<uint64_t>(ahi) << 32) | alo;
  const uint64_t dif = ((static_cast<uint64_t>(ohi) << 32) | olo) ^ all_and;  // nonzero: n > k >= 1 distinct
  int hi_bit = 63 - __clzll(static_cast<long long>(dif));
  uint64_t mask = hi_bit >= 63 ? 0ull : ~0ull << (hi_bit + 1);
  uint64_t prefix = all_and & mask;
  int need = k;
  for (;;) {
    const int width = hi_bit + 1 < 8 ? hi_bit + 1 : 8;
    const int shift = hi_bit + 1 - width;
    const uint32_t dm = (1u << width) - 1u;
    *reinterpret_cast<uint4*>(wh + 4 * lane) = make_uint4(0u, 0u, 0u, 0u);
    wave_sync();
    for (int e = lane; e < n; e += 64) {
      const uint64_t v = src[e];
      if ((v & mask) == prefix) atomicAdd(&wh[(v >> shift) & dm], 1u);
    }
    wave_sync();
    const uint4 h = *reinterpret_cast<const uint4*>(wh + 4 * lane);
    const int sum = static_cast<int>(h.x + h.y + h.z + h.w);
    const int incl = wave_incl_add(sum);
    const int excl = incl - sum;
    const bool hit = excl < need && need <= incl;
    int dd = 0, before = 0, cnt = 0;
    if (hit) {
      const int hb[4] = {static_cast<int>(h.x), static_cast<int>(h.y), static_cast<int>(h.z), static_cast<int>(h.w)};
      int c = excl;
      for (int q = 0; q < 4; ++q) {
        if (c + hb[q] >= need) {
          dd = 4 * lane + q;
          before = c;
          cnt = hb[q];
          break;
        }
        c += hb[q];
      }
    }
    const int hl = __ffsll(static_cast<long long>(__ballot(hit))) - 1;
    dd = __builtin_amdgcn_readlane(dd, hl);
    before = __builtin_amdgcn_readlane(before, hl);
    cnt = __builtin_amdgcn_readlane(cnt, hl);
    prefix |= static_cast<uint64_t>(dd) << shift;
    mask |= static_cast<uint64_t>(dm) << shift;
    need -= before;
    hi_bit = shift - 1;
    // distinct entries: once every bit is fixed the bucket holds one entry
    if (cnt == need || shift == 0) break;
    wave_sync();
  }
  const uint64_t top = prefix | ~mask;
  int got = 0;
  for (int e0 = 0; e0 < n; e0 += 64) {
    const int e = e0 + lane;
    const uint64_t v = e < n ? src[e] : ~0ull;
    const bool take = e < n && v <= top;
    const uint64_t b = __ballot(take);
    if (take) out[got + __popcll(b & ((1ull << lane) - 1ull))] = v;
    got += __popcll(b);
  }
  return got;
}

}  // namespace aknn
